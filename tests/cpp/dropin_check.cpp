// A C++ caller of the reference's dense surface (include/BEM3dLaplace.h,
// include/BEM3dElasticity.h), linked against libeasybem_hip.so as a
// reference caller would link the reference library.
//
//   dropin_check <input> <output>
//
// input (native endianness): int32 np, ne, order, npts, elastic; double nu;
// double nodes[3 np]; int32 elements[3 ne] (1-based); double points[3 npts];
// double neumann[ne]; double dirichlet[np].
// output: the doubles of V, K, D, M (getLaplace3d), M (getLaplace3dMass),
// C1, C2, C3 (getLaplace3dCurl), result (evaluateLaplace3d), then, when
// elastic, V, K, D, M of getElasticity3d, each in the layout the function
// returns, back to back.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "BEM3dElasticity.h"
#include "BEM3dLaplace.h"

namespace {
template <typename T>
bool get(std::FILE *f, T *p, size_t n)
{
  return std::fread(p, sizeof(T), n, f) == n;
}
void put(std::FILE *f, const double *p, size_t n) { std::fwrite(p, sizeof(double), n, f); }
}  // namespace

int main(int argc, char **argv)
{
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <input> <output>\n", argv[0]);
    return 2;
  }
  std::FILE *in = std::fopen(argv[1], "rb");
  if (!in) return 2;
  int32_t h[5];
  double nu = 0.0;
  if (!get(in, h, 5) || !get(in, &nu, 1)) return 2;
  const int np = h[0], ne = h[1], order = h[2], npts = h[3], elastic = h[4];
  std::vector<double> nodes(3 * (size_t)np), pts(3 * (size_t)npts), tn(ne), ud(np);
  std::vector<int> el(3 * (size_t)ne);
  if (!get(in, nodes.data(), nodes.size()) || !get(in, el.data(), el.size()) ||
      !get(in, pts.data(), pts.size()) || !get(in, tn.data(), tn.size()) ||
      !get(in, ud.data(), ud.size()))
    return 2;
  std::fclose(in);
  std::FILE *out = std::fopen(argv[2], "wb");
  if (!out) return 2;
  const size_t sne = ne, snp = np;

  double *V = nullptr, *K = nullptr, *D = nullptr, *M = nullptr;
  getLaplace3d(np, nodes.data(), ne, el.data(), V, K, D, M, order);
  put(out, V, sne * sne);
  put(out, K, sne * snp);
  put(out, D, snp * snp);
  put(out, M, sne * snp);
  delete[] V;
  delete[] K;
  delete[] D;
  delete[] M;

  double *Mm = nullptr;
  getLaplace3dMass(np, nodes.data(), ne, el.data(), Mm);
  put(out, Mm, sne * snp);
  delete[] Mm;

  double *C1 = nullptr, *C2 = nullptr, *C3 = nullptr;
  getLaplace3dCurl(np, nodes.data(), ne, el.data(), C1, C2, C3);
  for (double *C : {C1, C2, C3}) {
    put(out, C, sne * snp);
    delete[] C;
  }

  double *res = nullptr;
  evaluateLaplace3d(nodes.data(), ne, el.data(), npts, pts.data(), tn.data(), ud.data(), res,
                    order);
  put(out, res, npts);
  delete[] res;

  if (elastic) {
    getElasticity3d(np, nodes.data(), ne, el.data(), nu, V, K, D, M, order);
    put(out, V, 9 * sne * sne);
    put(out, K, 9 * sne * snp);
    put(out, D, 9 * snp * snp);
    put(out, M, 9 * sne * snp);
    delete[] V;
    delete[] K;
    delete[] D;
    delete[] M;
  }
  std::fclose(out);
  return 0;
}
