// Host-side mesh preparation for the layer-potential operators: quadrature
// rules, reduced Sauter-Schwab weight tables, mesh validation, touching-pair
// classification and near-field CSR patterns.  All of it runs once per mesh
// (or once per operator) on the CPU; nothing here is on the per-apply path.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace ebh {

// ---------------------------------------------------------------- quadrature
// Gauss-Legendre on [0,1]; sizes indexed like the reference table
// (src/quadrature.cpp:286-287).
int line_rule_size(int idx);
bool gauss01(int n, std::vector<double> &p, std::vector<double> &w);
bool line_rule(int idx, std::vector<double> &p, std::vector<double> &w);
// Symmetric triangle rules, index d-1 of polynomial degree d, weights summing
// to 1/2 (src/quadrature.cpp:1047-1048), indices 0..19 (1 to 79 points).
constexpr int kMaxTriPoints = 80;
bool tri_rule(int idx, std::vector<double> &x, std::vector<double> &y,
              std::vector<double> &w);
// Reorders a rule whose points are the centroid or lie on the medians
// (barycentric (a, a, 1-2a) and permutations) into: the centroid first, if
// present, then each orbit with its special vertex k = 0, 1, 2.  Point q is
// then C + c[q] (P_k - C), C the centroid and P_k the k-th vertex, with
// k = (q - has_centroid) % 3.  Returns false (rule untouched) for any other
// rule.
bool ray_order(int nt, double *x, double *y, double *w, double *c);

// Reduced Sauter-Schwab weight tables.  cls: 1 vertex-adjacent, 2
// edge-adjacent, 3 coincident.  kdeg: homogeneity degree of the kernel
// (1 single layer / Kelvin, 2 double layer).  Layout
// W[((m * nout) + q) * 9 + 3k + l], k test / l trial shape function in the
// pair's local vertex order.
int ss_nregions(int cls);
int ss_nouter(int cls, int nq);
void ss_table(int cls, int kdeg, int nq, std::vector<double> &W);

// ---------------------------------------------------------------- topology
enum PairClass : int { kDisjoint = 0, kVertex = 1, kEdge = 2, kCoincident = 3 };

struct NearPair {
  int i, j;       // test / trial triangle (0-based)
  int cls;        // PairClass
  int oi[3];      // global 0-based nodes of i in the pair's local order
  int oj[3];      // same for j
};

// row blocks of the fused far reduction and near-field pass
constexpr int kFinPos = 32, kFinRows = 32, kFinMax = 1024;

struct Topology {
  int np = 0, ne = 0;
  std::vector<double> nodes;            // np*3
  std::vector<int> tri;                 // ne*3, 0-based
  std::vector<int> node_tri_ptr, node_tri;    // node -> triangles (CSR)
  std::vector<int> node_tri_k;                // local index of node in tri
  // Blocks of node rows for the fused far reduction and near-field pass
  // (fin_blocks): block b is rows fin_blk[b] .. fin_blk[b+1]-1; its
  // incidences ntp[fin_blk[b]] .. ntp[fin_blk[b+1]]-1 sit at the padded
  // positions fin_pos[b] .. of a partial row, fin_pos[b] a multiple of
  // kFinPos; tri_pos[k * ne + t] is the padded position of (t, k), and
  // fin_pos.back() the padded row length
  std::vector<int> fin_blk, fin_pos, tri_pos;
  int fin_max = 0;
  std::vector<int> touch_ptr, touch;          // triangle -> touching triangles
                                              // (sorted ascending, incl. self)
  std::vector<NearPair> pairs;                // all touching pairs, (i, j) order
  int64_t n_class[4] = {0, 0, 0, 0};
  int max_touch = 0;
};

// Row blocks of the fused pass over node rows with incidence row pointers
// ntp (np + 1 entries): at most kFinRows rows and, unless one row alone has
// more, kFinPos incidences per block, greedily; blk gets the nblk + 1 row
// bounds, base the nblk + 1 padded first positions (multiples of kFinPos).
// Returns the most incidences of a block.
int fin_blocks(const std::vector<int> &ntp, std::vector<int> &blk, std::vector<int> &base);

// Validates the mesh (index range, non-degenerate triangles, every edge
// shared by exactly two triangles with opposite orientation) and builds the
// topology.  Returns an empty string on success, else the error.
std::string build_topology(int np, const double *nodes, int ne,
                           const int *elems_1based, Topology &T);

// Surface curls of the P1 hat functions, restating getLaplace3dCurl
// (src/BEM3dLaplace.cpp:489-558): curl[(3i + c) 3 + k] is Cartesian
// component c of curl_Gamma phi_k on triangle i, k its local vertex.
void surface_curls(const Topology &T, std::vector<double> &curl);

// Near-field CSR pattern for a (test basis, trial basis) combination with
// contribution lists: entry e sums pairbuf values listed in
// contrib[contrib_ptr[e] .. contrib_ptr[e+1]); a contribution is
// pair * 16 + code, code 0..8 = slot 3k+l, 9+l = sum over k (P0 test),
// 12+k = sum over l (P0 trial), 15 = sum of all nine (P0 x P0).
struct NearCSR {
  int nrows = 0, ncols = 0;
  std::vector<int> row_ptr, col;
  std::vector<int> contrib_ptr;
  std::vector<int64_t> contrib;
};
void build_near_csr(const Topology &T, int test_basis, int trial_basis,
                    NearCSR &out);

// Pass plan of the pair-symmetric multi-RHS schedule (far_field.hip,
// k_mrhs_tiles / k_mrhs_tile_mfma) for ne triangles in groups of 64 (ng
// groups).  Tiles are the group pairs (I, J), J >= I, row by row: tile (I,
// J) has index rowoff[I] + J - I.  Pass p covers tile rows rows[p] ..
// rows[p+1]-1 (blocks of `seg` rows, taken greedily while the pass's tiles
// at tile_bytes and items at item_bytes stay within `limit`, one block at
// least), its tiles from t0[p] on.  Its items (K, kind, M0, M1), 4 ints
// each, are item[4 ioff[p]] .. item[4 ioff[p+1]] - 1: for each output group
// K >= rows[p] the column items (kind 1: tiles (M, K), rows[p] <= M <
// min(rows[p+1], K)) in segments of at most `seg` on the `seg`-row grid,
// then, for K in the pass, the row items (kind 0: tiles (K, M), K <= M < ng)
// in segments of at most `seg`; kptr[koff[p] + K - rows[p]] is K's first
// item relative to ioff[p] (ng - rows[p] + 1 entries).  So every output's
// items run through M = 0 .. ng-1 in the same segments, in the same order,
// whatever the passes.  order[ioff[p] ..]: the pass's items (indices
// relative to ioff[p]) in launch order, longer items first, ties in list
// order.  max_tiles / max_items: the largest pass's counts.
struct MtPlan {
  int ng = 0;
  std::vector<int> rowoff, rows, t0, ioff, koff, item, kptr, order;
  size_t max_tiles = 0, max_items = 0;
};
void mt_plan(int ne, double tile_bytes, double item_bytes, double limit, int seg,
             MtPlan &p);

}  // namespace ebh
