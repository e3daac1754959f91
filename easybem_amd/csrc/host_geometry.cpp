// Host-side mesh preparation: quadrature, Sauter-Schwab tables, topology,
// near-field CSR patterns.  See host_geometry.h.
#include "host_geometry.h"

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>

namespace ebh {

// ---------------------------------------------------------------- quadrature
static const int kLineSizes[19] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12,
                                   16, 20, 24, 32, 40, 48, 64};

int line_rule_size(int idx)
{
  return (idx >= 0 && idx < 19) ? kLineSizes[idx] : -1;
}

bool gauss01(int n, std::vector<double> &p, std::vector<double> &w)
{
  if (n < 1 || n > 128) return false;
  p.assign(n, 0.0);
  w.assign(n, 0.0);
  auto legendre = [n](long double x, long double &pn, long double &pn1) {
    long double a = 1.0L, b = x;
    for (int k = 2; k <= n; k++) {
      long double c = ((2 * k - 1) * x * b - (k - 1) * a) / k;
      a = b;
      b = c;
    }
    if (n == 1) { a = 1.0L; b = x; }
    pn = b;
    pn1 = a;
  };
  for (int i = 0; i < (n + 1) / 2; i++) {
    long double x = cosl(3.14159265358979323846264338L * (i + 0.75L) /
                         (n + 0.5L));
    long double pn, pn1, dp = 1.0L;
    for (int it = 0; it < 100; it++) {
      legendre(x, pn, pn1);
      dp = n * (x * pn - pn1) / (x * x - 1.0L);
      long double dx = pn / dp;
      x -= dx;
      if (fabsl(dx) < 1e-19L) break;
    }
    legendre(x, pn, pn1);
    dp = n * (x * pn - pn1) / (x * x - 1.0L);
    long double wt = 1.0L / ((1.0L - x * x) * dp * dp);
    p[i] = (double)((1.0L - x) / 2.0L);
    p[n - 1 - i] = (double)((1.0L + x) / 2.0L);
    w[i] = w[n - 1 - i] = (double)wt;
  }
  if (n % 2) p[n / 2] = 0.5;
  return true;
}

bool line_rule(int idx, std::vector<double> &p, std::vector<double> &w)
{
  int n = line_rule_size(idx);
  return n > 0 && gauss01(n, p, w);
}

namespace {
struct TriRuleBuilder {
  std::vector<double> &x, &y, &w;
  void one(double a, double b, double wt) { x.push_back(a); y.push_back(b); w.push_back(wt); }
  void orbit3(double a, double wt)  // (a,a), (a,1-2a), (1-2a,a)
  {
    double b = 1.0 - 2.0 * a;
    one(a, a, wt); one(a, b, wt); one(b, a, wt);
  }
  void orbit6(double a, double b, double wt)  // permutations of (a,b,1-a-b)
  {
    double c = 1.0 - a - b;
    one(a, b, wt); one(b, c, wt); one(c, a, wt);
    one(b, a, wt); one(c, b, wt); one(a, c, wt);
  }
};
}  // namespace

// Published symmetric rules: Strang & Fix (1, 3, 4 points), Radon's closed
// form (7 points, degree 5), Dunavant 1985 (6, 12, 13 points).
bool tri_rule(int idx, std::vector<double> &x, std::vector<double> &y,
              std::vector<double> &w)
{
  x.clear(); y.clear(); w.clear();
  TriRuleBuilder R{x, y, w};
  switch (idx) {
  case 0: R.one(1.0 / 3.0, 1.0 / 3.0, 0.5); return true;
  case 1: R.orbit3(1.0 / 6.0, 1.0 / 6.0); return true;
  case 2:
    R.one(1.0 / 3.0, 1.0 / 3.0, -27.0 / 96.0);
    R.orbit3(0.2, 25.0 / 96.0);
    return true;
  case 3:
    R.orbit3(0.445948490915965, 0.223381589678011 / 2);
    R.orbit3(0.091576213509771, 0.109951743655322 / 2);
    return true;
  case 4: {
    double s = std::sqrt(15.0);
    R.one(1.0 / 3.0, 1.0 / 3.0, 9.0 / 80.0);
    R.orbit3((6.0 - s) / 21.0, (155.0 - s) / 2400.0);
    R.orbit3((6.0 + s) / 21.0, (155.0 + s) / 2400.0);
    return true;
  }
  case 5:
    R.orbit3(0.249286745170910, 0.116786275726379 / 2);
    R.orbit3(0.063089014491502, 0.050844906370207 / 2);
    R.orbit6(0.053145049844817, 0.310352451033784, 0.082851075618374 / 2);
    return true;
  case 6:
    R.one(1.0 / 3.0, 1.0 / 3.0, -0.149570044467682 / 2);
    R.orbit3(0.260345966079040, 0.175615257433208 / 2);
    R.orbit3(0.065130102902216, 0.053347235608838 / 2);
    R.orbit6(0.048690315425316, 0.312865496004874, 0.077113760890257 / 2);
    return true;
  default: return false;
  }
}

bool ray_order(int nt, double *x, double *y, double *w, double *c)
{
  const double tol = 1e-13;
  int cen = -1;
  std::vector<int> kk(nt, -1);
  std::vector<double> cc(nt, 0.0);
  for (int q = 0; q < nt; q++) {
    const double mu[3] = {1.0 - x[q] - y[q] - 1.0 / 3.0, x[q] - 1.0 / 3.0,
                          y[q] - 1.0 / 3.0};
    if (std::fabs(mu[0]) < tol && std::fabs(mu[1]) < tol && std::fabs(mu[2]) < tol) {
      if (cen >= 0) return false;
      cen = q;
      continue;
    }
    for (int k = 0; k < 3; k++) {
      const int i = (k + 1) % 3, j = (k + 2) % 3;
      if (std::fabs(mu[i] - mu[j]) < tol) { kk[q] = k; cc[q] = mu[k] - mu[i]; }
    }
    if (kk[q] < 0) return false;
  }
  // orbits in order of first appearance; each needs k = 0, 1, 2 once
  std::vector<int> order;
  if (cen >= 0) order.push_back(cen);
  std::vector<bool> used(nt, false);
  if (cen >= 0) used[cen] = true;
  for (int q = 0; q < nt; q++) {
    if (used[q]) continue;
    int mem[3] = {-1, -1, -1};
    for (int p = q; p < nt; p++)
      if (!used[p] && kk[p] >= 0 && std::fabs(cc[p] - cc[q]) < 1e-12) {
        if (mem[kk[p]] >= 0) return false;
        mem[kk[p]] = p;
      }
    for (int k = 0; k < 3; k++) {
      if (mem[k] < 0) return false;
      used[mem[k]] = true;
      order.push_back(mem[k]);
      // one radius per orbit (the members' agree to rounding)
      cc[mem[k]] = cc[q];
    }
  }
  std::vector<double> ox(nt), oy(nt), ow(nt);
  for (int q = 0; q < nt; q++) {
    ox[q] = x[order[q]];
    oy[q] = y[order[q]];
    ow[q] = w[order[q]];
    c[q] = order[q] == cen ? 0.0 : cc[order[q]];
  }
  for (int q = 0; q < nt; q++) { x[q] = ox[q]; y[q] = oy[q]; w[q] = ow[q]; }
  return true;
}

// ------------------------------------------------------- Sauter-Schwab maps
int ss_nregions(int cls) { return cls == 1 ? 2 : cls == 2 ? 5 : 3; }
int ss_nouter(int cls, int nq)
{
  return cls == 1 ? nq * nq * nq : cls == 2 ? nq * nq : nq;
}

// Maps (inner, outer) -> (s, t) on the reference triangle {0<=s2<=s1<=1},
// Jacobian and homogeneity factor h (x - y = h * d_m).  Vertex- and
// edge-adjacent maps: Sauter & Schwab, Boundary Element Methods, sec. 5.2.1,
// regions ordered so that d_m matches the reduced difference vectors of
// src/BEM3dLaplace.cpp:76-99,133-142.  Coincident: the difference z = s - t
// runs along six rays z = rho*zhat_m(eta) of the hexagon T - T; for fixed z,
// s runs over T n (T+z), a triangle of side 1 - rho.
static void ss_map(int cls, int m, const double *in, const double *out,
                   double *s, double *t, double *jac, double *h)
{
  if (cls == 1) {
    double xi = in[0], e1 = out[0], e2 = out[1], e3 = out[2];
    double a0 = xi, a1 = xi * e1, b0 = xi * e2, b1 = xi * e2 * e3;
    if (m == 0) { s[0] = a0; s[1] = a1; t[0] = b0; t[1] = b1; }
    else { s[0] = b0; s[1] = b1; t[0] = a0; t[1] = a1; }
    *jac = xi * xi * xi * e2;
    *h = xi;
  } else if (cls == 2) {
    double xi = in[0], e1 = in[1], e2 = out[0], e3 = out[1];
    switch (m) {
    case 0:
      s[0] = xi; s[1] = xi * e1;
      t[0] = xi * (1 - e1 * e2 * e3); t[1] = xi * e1 * e2 * (1 - e3); break;
    case 1:
      s[0] = xi * (1 - e1 * e2); s[1] = xi * e1 * (1 - e2);
      t[0] = xi; t[1] = xi * e1 * e2 * e3; break;
    case 2:
      s[0] = xi * (1 - e1 * e2 * e3); s[1] = xi * e1 * e2 * (1 - e3);
      t[0] = xi; t[1] = xi * e1; break;
    case 3:
      s[0] = xi * (1 - e1 * e2 * e3); s[1] = xi * e1 * (1 - e2 * e3);
      t[0] = xi; t[1] = xi * e1 * e2; break;
    default:
      s[0] = xi; s[1] = xi * e1 * e3;
      t[0] = xi * (1 - e1 * e2); t[1] = xi * e1 * (1 - e2); break;
    }
    *jac = xi * xi * xi * e1 * e1 * (m < 4 ? e2 : 1.0);
    *h = xi * e1;
  } else {
    double rho = in[0], l1 = in[1], l2 = in[2], eta = out[0], zh0, zh1;
    switch (m % 3) {
    case 0: zh0 = eta; zh1 = 1.0; break;
    case 1: zh0 = 1.0; zh1 = eta; break;
    default: zh0 = eta; zh1 = eta - 1.0; break;
    }
    if (m >= 3) { zh0 = -zh0; zh1 = -zh1; }
    double z1 = rho * zh0, z2 = rho * zh1;
    double a = std::max(z2, 0.0), b = 1.0 + std::min(z1, 0.0);
    double c = std::max(z1 - z2, 0.0);
    double V1x = a + c, V1y = a, V2x = b, V2y = a, V3x = b, V3y = b - c;
    s[0] = V1x + l1 * (V2x - V1x) + l2 * (V3x - V1x);
    s[1] = V1y + l1 * (V2y - V1y) + l2 * (V3y - V1y);
    t[0] = s[0] - z1;
    t[1] = s[1] - z2;
    double side = b - a - c;
    *jac = rho * side * side;
    *h = rho;
  }
}

// Exact inner integration: the inner integrands are polynomials of degree
// <= 5 per variable, integrated by 4-point Gauss-Legendre and (coincident
// simplex) the degree-4 six-point rule.  The coincident table folds the
// opposite rays m and m+3, valid for kernels even in x - y.
void ss_table(int cls, int kdeg, int nq, std::vector<double> &W)
{
  std::vector<double> L, Lw, G, Gw, Sx, Sy, Sw;
  gauss01(nq, L, Lw);
  gauss01(4, G, Gw);
  tri_rule(3, Sx, Sy, Sw);
  const int nreg = ss_nregions(cls), nout = ss_nouter(cls, nq);
  W.assign((size_t)9 * nreg * nout, 0.0);
  auto add = [&](double *Wm, double c, const double *s, const double *t) {
    double fs[3] = {1.0 - s[0], s[0] - s[1], s[1]};
    double ft[3] = {1.0 - t[0], t[0] - t[1], t[1]};
    for (int k = 0; k < 3; k++)
      for (int l = 0; l < 3; l++) Wm[3 * k + l] += c * fs[k] * ft[l];
  };
  for (int q = 0; q < nout; q++) {
    double out[3] = {0, 0, 0};
    if (cls == 1) {
      out[0] = L[q / (nq * nq)]; out[1] = L[(q / nq) % nq]; out[2] = L[q % nq];
    } else if (cls == 2) {
      out[0] = L[q / nq]; out[1] = L[q % nq];
    } else {
      out[0] = L[q];
    }
    const int mreg = cls == 3 ? 6 : nreg;
    for (int m = 0; m < mreg; m++) {
      double *Wm = W.data() + ((size_t)(m % nreg) * nout + q) * 9;
      double s[2], t[2], jac, h;
      if (cls == 1) {
        for (size_t g = 0; g < G.size(); g++) {
          double in[1] = {G[g]};
          ss_map(cls, m, in, out, s, t, &jac, &h);
          add(Wm, Gw[g] * jac / std::pow(h, kdeg), s, t);
        }
      } else if (cls == 2) {
        for (size_t g = 0; g < G.size(); g++)
          for (size_t g2 = 0; g2 < G.size(); g2++) {
            double in[2] = {G[g], G[g2]};
            ss_map(cls, m, in, out, s, t, &jac, &h);
            add(Wm, Gw[g] * Gw[g2] * jac / std::pow(h, kdeg), s, t);
          }
      } else {
        for (size_t g = 0; g < G.size(); g++)
          for (size_t r = 0; r < Sw.size(); r++) {
            double in[3] = {G[g], Sx[r], Sy[r]};
            ss_map(cls, m, in, out, s, t, &jac, &h);
            add(Wm, Gw[g] * Sw[r] * jac / std::pow(h, kdeg), s, t);
          }
      }
    }
  }
}

// ---------------------------------------------------------------- topology
std::string build_topology(int np, const double *nodes, int ne,
                           const int *elems, Topology &T)
{
  if (np < 3 || ne < 4) return "mesh too small (need >= 3 nodes, >= 4 triangles)";
  T = Topology();
  T.np = np;
  T.ne = ne;
  T.nodes.assign(nodes, nodes + (size_t)3 * np);
  for (size_t k = 0; k < T.nodes.size(); k++)
    if (!std::isfinite(T.nodes[k])) return "non-finite node coordinate";
  T.tri.resize((size_t)3 * ne);
  for (int i = 0; i < ne; i++) {
    for (int k = 0; k < 3; k++) {
      int n = elems[3 * i + k];
      if (n < 1 || n > np)
        return "element " + std::to_string(i) + " references node " +
               std::to_string(n) + " outside 1.." + std::to_string(np);
      T.tri[3 * i + k] = n - 1;
    }
    const int *f = &T.tri[3 * i];
    if (f[0] == f[1] || f[1] == f[2] || f[0] == f[2])
      return "element " + std::to_string(i) + " repeats a node";
    const double *a = &T.nodes[3 * f[0]], *b = &T.nodes[3 * f[1]],
                 *c = &T.nodes[3 * f[2]];
    double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double n0 = u[1] * v[2] - u[2] * v[1], n1 = u[2] * v[0] - u[0] * v[2],
           n2 = u[0] * v[1] - u[1] * v[0];
    double J = std::sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    double L2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2] + v[0] * v[0] +
                v[1] * v[1] + v[2] * v[2];
    if (!(J > 1e-12 * L2)) return "element " + std::to_string(i) + " is degenerate";
  }
  const int nthreads =
      (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  // runs fn(tid) on host threads; item r of a loop goes to thread r mod
  // nthreads, and every item below is independent of the others
  auto on_threads = [&](auto fn) {
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; t++) pool.emplace_back(fn, t);
    fn(0);
    for (auto &th : pool) th.join();
  };
  // node -> triangles
  T.node_tri_ptr.assign(np + 1, 0);
  for (int i = 0; i < 3 * ne; i++) T.node_tri_ptr[T.tri[i] + 1]++;
  for (int n = 0; n < np; n++) T.node_tri_ptr[n + 1] += T.node_tri_ptr[n];
  T.node_tri.resize((size_t)3 * ne);
  T.node_tri_k.resize((size_t)3 * ne);
  {
    std::vector<int> fill(T.node_tri_ptr.begin(), T.node_tri_ptr.end() - 1);
    for (int i = 0; i < ne; i++)
      for (int k = 0; k < 3; k++) {
        int n = T.tri[3 * i + k];
        T.node_tri[fill[n]] = i;
        T.node_tri_k[fill[n]] = k;
        fill[n]++;
      }
  }
  // closed, consistently oriented 2-manifold: each undirected edge appears
  // exactly twice, once in each direction.  Edge {a, b}, a < b, is checked
  // at node a: around a, a -> b is the edge to the next vertex of a
  // triangle and b -> a the edge from the previous one.  The failing edge
  // of smallest (a, b) is reported.
  {
    struct Bad { int64_t key = INT64_MAX; int count = 0; bool orient = false; };
    std::vector<Bad> bad(nthreads);
    on_threads([&](int tid) {
      std::vector<int> nb, out, in;
      for (int a = tid; a < np; a += nthreads) {
        nb.clear();
        out.clear();
        in.clear();
        auto slot = [&](int b) {
          for (size_t q = 0; q < nb.size(); q++)
            if (nb[q] == b) return (int)q;
          nb.push_back(b);
          out.push_back(0);
          in.push_back(0);
          return (int)nb.size() - 1;
        };
        for (int p = T.node_tri_ptr[a]; p < T.node_tri_ptr[a + 1]; p++) {
          const int t = T.node_tri[p], k = T.node_tri_k[p];
          out[slot(T.tri[3 * t + (k + 1) % 3])]++;
          in[slot(T.tri[3 * t + (k + 2) % 3])]++;
        }
        for (size_t q = 0; q < nb.size(); q++) {
          if (nb[q] < a) continue;
          const int count = out[q] + in[q];
          const bool fail = count != 2 || out[q] != in[q];
          const int64_t key = (int64_t)a * np + nb[q];
          if (fail && key < bad[tid].key) bad[tid] = Bad{key, count, count == 2};
        }
      }
    });
    Bad worst;
    for (const Bad &b : bad)
      if (b.key < worst.key) worst = b;
    if (worst.key != INT64_MAX) {
      const std::string edge = "edge (" + std::to_string(worst.key / np + 1) + "," +
                               std::to_string(worst.key % np + 1) + ")";
      if (!worst.orient)
        return edge + " is shared by " + std::to_string(worst.count) +
               " triangles: the mesh must be a closed 2-manifold";
      return edge + " has the same direction in both triangles: inconsistent "
                    "orientation";
    }
  }
  for (int n = 0; n < np; n++)
    if (T.node_tri_ptr[n + 1] == T.node_tri_ptr[n])
      return "node " + std::to_string(n + 1) + " belongs to no triangle";
  // touching lists and classified pairs, in (i, j) order: the touching
  // triangles of i (its vertices' triangles, sorted, self included) are
  // counted, then filled at their offsets
  T.touch_ptr.assign(ne + 1, 0);
  auto touching = [&](int i, std::vector<int> &cand) {
    cand.clear();
    for (int k = 0; k < 3; k++) {
      int n = T.tri[3 * i + k];
      for (int p = T.node_tri_ptr[n]; p < T.node_tri_ptr[n + 1]; p++)
        cand.push_back(T.node_tri[p]);
    }
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
  };
  on_threads([&](int tid) {
    std::vector<int> cand;
    for (int i = tid; i < ne; i += nthreads) {
      touching(i, cand);
      T.touch_ptr[i + 1] = (int)cand.size();
    }
  });
  for (int i = 0; i < ne; i++) {
    T.max_touch = std::max(T.max_touch, T.touch_ptr[i + 1]);
    T.touch_ptr[i + 1] += T.touch_ptr[i];
  }
  T.touch.resize(T.touch_ptr[ne]);
  T.pairs.resize(T.touch_ptr[ne]);
  std::vector<std::array<int64_t, 4>> ncls(nthreads, std::array<int64_t, 4>{0, 0, 0, 0});
  std::vector<int> dup(nthreads, INT_MAX), dupj(nthreads, 0);
  on_threads([&](int tid) {
    std::vector<int> cand;
    for (int i = tid; i < ne; i += nthreads) {
      touching(i, cand);
      const int *fi = &T.tri[3 * i];
      int e = T.touch_ptr[i];
      for (int j : cand) {
        T.touch[e] = j;
        const int *fj = &T.tri[3 * j];
        NearPair P;
        P.i = i;
        P.j = j;
        if (i == j) {
          P.cls = kCoincident;
          for (int k = 0; k < 3; k++) P.oi[k] = P.oj[k] = fi[k];
        } else {
          // shared nodes scanned in j's order (src/BEM3dLaplace.cpp:692-709)
          int ci[3], cj[3], cnt = 0;
          for (int k = 0; k < 3; k++)
            for (int l = 0; l < 3; l++)
              if (fj[k] == fi[l]) { ci[cnt] = l; cj[cnt] = k; cnt++; break; }
          if (cnt == 1) {
            P.cls = kVertex;
            for (int k = 0; k < 3; k++) {
              P.oi[k] = fi[(ci[0] + k) % 3];
              P.oj[k] = fj[(cj[0] + k) % 3];
            }
          } else if (cnt == 2) {
            // orient so the shared edge runs cj0 -> cj0+1 in j
            if (cj[0] == 0 && cj[1] == 2) {
              std::swap(cj[0], cj[1]);
              std::swap(ci[0], ci[1]);
            }
            int remi = 3 - ci[0] - ci[1], remj = 3 - cj[0] - cj[1];
            P.cls = kEdge;
            P.oi[0] = fi[ci[0]]; P.oi[1] = fi[ci[1]]; P.oi[2] = fi[remi];
            P.oj[0] = fj[cj[0]]; P.oj[1] = fj[cj[1]]; P.oj[2] = fj[remj];
          } else {
            // reported for the smallest i (then j) after the loop
            if (i < dup[tid]) { dup[tid] = i; dupj[tid] = j; }
            P.cls = kCoincident;
          }
        }
        ncls[tid][P.cls]++;
        T.pairs[e] = P;
        e++;
      }
    }
  });
  int di = INT_MAX, dj = 0;
  for (int t = 0; t < nthreads; t++)
    if (dup[t] < di) { di = dup[t]; dj = dupj[t]; }
  if (di != INT_MAX)
    return "triangles " + std::to_string(di) + " and " + std::to_string(dj) +
           " share all three nodes";
  for (int t = 0; t < nthreads; t++)
    for (int c = 0; c < 4; c++) T.n_class[c] += ncls[t][c];
  T.n_class[kDisjoint] = (int64_t)ne * ne - (int64_t)T.pairs.size();
  return "";
}

void build_near_csr(const Topology &T, int test_basis, int trial_basis,
                    NearCSR &out)
{
  out = NearCSR();
  out.nrows = test_basis ? T.np : T.ne;
  out.ncols = trial_basis ? T.np : T.ne;
  struct Item { int col; int64_t c; };
  const int nl = trial_basis ? 3 : 1;
  // The pairs of test triangle i are T.pairs[touch_ptr[i] .. touch_ptr[i+1]),
  // so every row's contributions can be gathered on its own: row a of a P1
  // test basis takes the pairs of the triangles around node a (local test
  // node k with oi[k] == a), a P0 row those of its triangle.  Rows run on
  // host threads (row r on thread r mod nthreads); a row's items are sorted
  // by (column, contribution), which is unique per item, so the result does
  // not depend on the gathering order or the split.
  auto for_items = [&](int r, auto &&fn) {
    auto pair_items = [&](int p, int k) {
      const NearPair &P = T.pairs[p];
      for (int l = 0; l < nl; l++) {
        const int col = trial_basis ? P.oj[l] : P.j;
        int code;
        if (test_basis && trial_basis) code = 3 * k + l;
        else if (!test_basis && trial_basis) code = 9 + l;
        else if (test_basis && !trial_basis) code = 12 + k;
        else code = 15;
        fn(col, (int64_t)p * 16 + code);
      }
    };
    if (test_basis) {
      for (int q = T.node_tri_ptr[r]; q < T.node_tri_ptr[r + 1]; q++) {
        const int i = T.node_tri[q];
        for (int p = T.touch_ptr[i]; p < T.touch_ptr[i + 1]; p++) {
          const NearPair &P = T.pairs[p];
          const int k = P.oi[0] == r ? 0 : P.oi[1] == r ? 1 : 2;
          pair_items(p, k);
        }
      }
    } else {
      for (int p = T.touch_ptr[r]; p < T.touch_ptr[r + 1]; p++) pair_items(p, 0);
    }
  };
  const int nthreads =
      (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  auto on_threads = [&](auto fn) {
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; t++) pool.emplace_back(fn, t);
    fn(0);
    for (auto &th : pool) th.join();
  };
  // pass 1: items and distinct columns per row (a visit stamp per column)
  std::vector<int> nitem(out.nrows), ncol(out.nrows);
  on_threads([&](int tid) {
    std::vector<int> stamp(out.ncols, -1);
    for (int r = tid; r < out.nrows; r += nthreads) {
      int ni = 0, nc = 0;
      for_items(r, [&](int col, int64_t) {
        ni++;
        if (stamp[col] != r) { stamp[col] = r; nc++; }
      });
      nitem[r] = ni;
      ncol[r] = nc;
    }
  });
  std::vector<int> cnt(out.nrows + 1, 0);
  out.row_ptr.assign(out.nrows + 1, 0);
  for (int r = 0; r < out.nrows; r++) {
    cnt[r + 1] = cnt[r] + nitem[r];
    out.row_ptr[r + 1] = out.row_ptr[r] + ncol[r];
  }
  const int nnz = out.row_ptr[out.nrows];
  out.col.resize(nnz);
  out.contrib_ptr.resize(nnz + 1);
  out.contrib.resize(cnt[out.nrows]);
  // pass 2: gather, sort and emit every row; entry starts are the first
  // item of each column run
  on_threads([&](int tid) {
    std::vector<Item> items;
    for (int r = tid; r < out.nrows; r += nthreads) {
      items.clear();
      for_items(r, [&](int col, int64_t c) { items.push_back(Item{col, c}); });
      std::sort(items.begin(), items.end(), [](const Item &x, const Item &y) {
        return x.col != y.col ? x.col < y.col : x.c < y.c;
      });
      int e = out.row_ptr[r];
      for (size_t i = 0; i < items.size(); i++) {
        out.contrib[cnt[r] + i] = items[i].c;
        if (i == 0 || items[i].col != items[i - 1].col) {
          out.col[e] = items[i].col;
          out.contrib_ptr[e] = cnt[r] + (int)i;
          e++;
        }
      }
    }
  });
  out.contrib_ptr[nnz] = cnt[out.nrows];
}

// The surface gradient of the hat functions solves the 2x2 metric system of
// the edges e1 = P1 - P0, e2 = P2 - P0; its rotation by the unit normal n,
// b = e x n per edge, gives the curl (src/BEM3dLaplace.cpp:508-547).
void surface_curls(const Topology &T, std::vector<double> &curl)
{
  curl.assign((size_t)9 * T.ne, 0.0);
  for (int i = 0; i < T.ne; i++) {
    const double *P[3];
    for (int k = 0; k < 3; k++) P[k] = &T.nodes[3 * T.tri[3 * i + k]];
    double e1[3], e2[3];
    for (int k = 0; k < 3; k++) { e1[k] = P[1][k] - P[0][k]; e2[k] = P[2][k] - P[0][k]; }
    double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                   e1[0] * e2[1] - e1[1] * e2[0]};
    const double J = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (double &c : n) c /= J;
    const double a11 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
    const double a22 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
    const double a12 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
    const double det = a11 * a22 - a12 * a12;
    const double b1[3] = {-e1[1] * n[2] + e1[2] * n[1], e1[0] * n[2] - e1[2] * n[0],
                          -e1[0] * n[1] + e1[1] * n[0]};
    const double b2[3] = {-e2[1] * n[2] + e2[2] * n[1], e2[0] * n[2] - e2[2] * n[0],
                          -e2[0] * n[1] + e2[1] * n[0]};
    for (int c = 0; c < 3; c++) {
      const double y1 = (b1[c] * a22 - b2[c] * a12) / det;
      const double y2 = (a11 * b2[c] - a12 * b1[c]) / det;
      double *o = &curl[(3 * (size_t)i + c) * 3];
      o[0] = -y1 - y2;
      o[1] = y1;
      o[2] = y2;
    }
  }
}

}  // namespace ebh
