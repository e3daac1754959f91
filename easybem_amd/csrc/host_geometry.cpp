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

// Dunavant's symmetric rules of degree 8 to 20 (D. A. Dunavant, High degree
// efficient symmetrical Gaussian quadrature rules for the triangle, Int. J.
// Numer. Meth. Eng. 21, 1985), the reference's triangle-rule indices 7..19
// (src/quadrature.cpp:1047-1048; 16 to 79 points), with the weights halved
// so that they sum to 1/2.  In orbit form: per rule the centroid weight (0:
// no centroid point), then kDunN3 orbits {a, w} of three points (a, a),
// (a, 1-2a), (1-2a, a) and kDunN6 orbits {a, b, w} of six points, the
// permutations of the barycentric coordinates (a, b, 1-a-b), back to back in
// kDunOrb.  The degree-20 orbit with a = 0.45513668195028 is Dunavant's:
// the reference lists 0.45513668695028 for it (a transcription error that
// leaves its three points off the orbit by 5e-9).
static const int kDunN3[13] = {3, 4, 2, 5, 5, 6, 6, 6, 7, 8, 9, 8, 10};
static const int kDunN6[13] = {1, 1, 3, 2, 3, 3, 4, 5, 5, 6, 7, 8, 8};
static const double kDunCen[13] = {
    0.072157803838893503, 0.048567898141399501, 0.045408995191376998,
    0, 0, 0.026260461700401,
    0, 0, 0.023437848713821,
    0.0167185996454015, 0.0154049699688235, 0.016453165694459499,
    0.016528527770811999,
};
static const double kDunOrb[] = {
    /* index 7 */
    0.45929258829272301, 0.047545817133642497, 0.17056930775175999,
    0.051608685267358997, 0.050547228317031012, 0.016229248811599001,
    0.263112829634638, 0.72849239295540402, 0.0136151570872175,
    /* index 8 */
    0.48968251919873801, 0.015667350113569501, 0.43708959149293702,
    0.038913770502387, 0.18820353561903302, 0.039823869463604999,
    0.044729513394453024, 0.012788837829349, 0.22196298916076601,
    0.74119859878449801, 0.021641769688644501,
    /* index 9 */
    0.485577633383657, 0.018362978878233498, 0.10948157548503701, 0.022660529717764,
    0.30793983876412101, 0.55035294182099903, 0.036378958422710002,
    0.246672560639903, 0.72832390459741103, 0.0141636212655285,
    0.066803251012200027, 0.92365593358749998, 0.0047108334818665,
    /* index 10 */
    0.534611048270758, 0.0004635031644805, 0.39896930296585498,
    0.038574767457406497, 0.203309900431282, 0.029661488690387001,
    0.11935091228258099, 0.018092270251709001, 0.032364948111275993,
    0.0068298655013389999, 0.35662064826129303, 0.59320121342821297,
    0.026168555981102001, 0.17148898030404203, 0.807489003159792,
    0.0103538298195705,
    /* index 11 */
    0.48821738977380502, 0.0128655332202275, 0.43972439229445998,
    0.021846272269019001, 0.27121038501211603, 0.031429112108942503,
    0.12757614554158597, 0.0173980564653545, 0.021317350453209982,
    0.0030831305257795001, 0.27571326968551402, 0.60894323577978804,
    0.020185778883190501, 0.28132558098993998, 0.69583608678780307,
    0.0111783866011515, 0.11625191590759698, 0.85801403354407302,
    0.0086581155543294999,
    /* index 12 */
    0.49504818493970498, 0.0056400726046649998, 0.46871663510957401,
    0.015711759181227001, 0.41452133680127701, 0.023536251252097001,
    0.22939957204283101, 0.023681793268177501, 0.11442449519633002,
    0.015583764522897, 0.024811391363459001, 0.0039878857325369998,
    0.26879499705876098, 0.63635117456166002, 0.018424201364366,
    0.29173006673428803, 0.69016915998690498, 0.0087007316519110006,
    0.12635738549166903, 0.851409537834241, 0.0077608934195224997,
    /* index 13 */
    0.48896391036217901, 0.0109417906847145, 0.417644719340454,
    0.016394176772062501, 0.27347752830883898, 0.025887052253646001,
    0.177205532412543, 0.021081294368496501, 0.061799883090872976,
    0.0072168498348885004, 0.019390961248700989, 0.0024617018012000002,
    0.17226668782135601, 0.77060855477499601, 0.012332876606282,
    0.33686145979634496, 0.57022229084668297, 0.019285755393530502,
    0.29837288213625801, 0.68698016780808802, 0.007218154056767, 0.118974497696957,
    0.87975717137017106, 0.0025051144192504999,
    /* index 14 */
    0.50697291685824297, 0.00095843782142449995, 0.43140635428302299,
    0.022124513635572499, 0.27769364484714398, 0.025593274359425999,
    0.12646489104125402, 0.011843867935343999, 0.070808385974685972,
    0.0066448878450105003, 0.018965170241073026, 0.002374458304096,
    0.26131137114008701, 0.60495446689329102, 0.019275036299796499,
    0.38804676709026897, 0.57558655551281401, 0.013607907160312,
    0.28571222004991603, 0.72446266307665497, 0.0010910386833985,
    0.21559966407228398, 0.74755646605183801, 0.010752659923865499,
    0.10357561657638598, 0.88396457409241602, 0.0038369713155245,
    /* index 15 */
    0.497380541948438, 0.0032029392892924998, 0.41346943854935203,
    0.020855148369693498, 0.470458599066991, 0.013445742125032, 0.240553749969521,
    0.021066261380824999, 0.147965794222573, 0.015000133421386499,
    0.075465187657474009, 0.0071000494625119998, 0.016596402623024975,
    0.0017912311756365, 0.296555596579887, 0.59986871117486096,
    0.016386573730313499, 0.337723063403079, 0.64219352494150495,
    0.0076491531242204999, 0.204748281642812, 0.79959272097132694,
    0.0011931220964195, 0.189358492130623, 0.768699721401368, 0.0095423963779494997,
    0.085283615682657021, 0.90039906408666104, 0.0034250272732710002,
    /* index 16 */
    0.49717054055677401, 0.0025467077202535001, 0.48217632262462501,
    0.0073354322638189999, 0.450239969020782, 0.012175439176836,
    0.40026623937739703, 0.015553775434484499, 0.25214126797095299,
    0.015628555609310001, 0.162047004658461, 0.0124078271698325,
    0.075875882260745986, 0.0070280365352785, 0.015654726967821997,
    0.0015973380868894999, 0.33431986736365804, 0.655493203809423,
    0.0040598276594965002, 0.29222153779694399, 0.57233759053202005,
    0.0134028711415815, 0.31957488542319001, 0.62600119028622803,
    0.0092299966054109998, 0.19070422419229199, 0.79642721497407098,
    0.0042384342671640002, 0.18048321164874598, 0.75235100593772897,
    0.0091463983850124993, 0.080711313679564023, 0.90462550409560794,
    0.0033328160020825001,
    /* index 17 */
    0.49334480863092101, 0.0045362183397019999, 0.46921059424195699,
    0.009380658469797, 0.43628139588700598, 0.0097205489927384998,
    0.39484617067341599, 0.013876974305405, 0.24979456880315698,
    0.016128112675728501, 0.16143219374384299, 0.012537016308461,
    0.076598227485370984, 0.0076359639859159998, 0.024252439353450006,
    0.0033969610114815002, 0.04314636721696502, -0.00111154936496,
    0.35891149494094399, 0.63265796885663605, 0.0031659570382029999,
    0.29440247675195697, 0.57441097151085496, 0.013628769024569,
    0.32501780164181404, 0.62477904679251206, 0.0088383928247324998,
    0.18473755966604599, 0.74893317652303704, 0.0091897423190350006,
    0.21879680001332102, 0.76920700542044296, 0.0040523664040959997,
    0.10117959713640801, 0.88396230227346706, 0.0038170645353625002,
    0.020874755282585977, 1.0143472600053629, 2.3093830397e-05,
    /* index 18 */
    0.48960998707300601, 0.0051653659456359999, 0.45453689269789299,
    0.011193623631508, 0.40141668064943103, 0.015133062934733999,
    0.25555165440309802, 0.015245483901099, 0.17707794215213002,
    0.012079606370820501, 0.11006105322795201, 0.0080254017934004995,
    0.05552862425184002, 0.004042290130892, 0.012621863777229014,
    0.0010396810137425, 0.39575478735694303, 0.60063379479464496,
    0.0019424384524905, 0.30792998388043602, 0.55760326158878404,
    0.012787080306011001, 0.26456694840652001, 0.72098702581736496,
    0.0044404517866689997, 0.35853935220595101, 0.59452706895587104,
    0.0080622733808655006, 0.15780740596859499, 0.83933147368083905,
    0.0012459709087455, 0.075050596975910988, 0.70108797892617303,
    0.0091214200594755008, 0.14242160111338298, 0.82293132406985703,
    0.0051292818680994997, 0.065494628082937989, 0.92434425262078401,
    0.001899964427651,
    /* index 19 */
    0.50095046435220003, 0.00043350959283149998, 0.48821295793472852,
    0.0058300263582240002, 0.45513668195028251, 0.0114384681782105,
    0.40199625931828953, 0.015224491336968999, 0.25589290975942147,
    0.0153124458626775, 0.17648825599510598, 0.0121840288384, 0.10417085533675852,
    0.0079987160160120002, 0.053068963840930017, 0.0038491509078010001,
    0.041618715196028999, -0.00031603024874399997, 0.011581921406822016,
    0.00087556715059649995, 0.34485577022900099, 0.60640264610616001,
    0.0082329195947880001, 0.37784326959485398, 0.61584261445654098,
    0.0024195167702425, 0.306635479062357, 0.55904800039029501, 0.012902453267325,
    0.24941936277474203, 0.73660674326286601, 0.0042355455272205002,
    0.21277572480280199, 0.711675142287434, 0.0091774570531399993,
    0.14696543605323897, 0.86140271715498695, 0.00035220233895399999,
    0.13772697882892299, 0.83558695791236293, 0.0050563424637309999,
    0.059696109149007026, 0.92975617155685297, 0.001786954692975,
};

// Published symmetric rules: Strang & Fix (1, 3, 4 points), Radon's closed
// form (7 points, degree 5), Dunavant 1985 (6, 12, 13 points; 16 to 79
// points from the table above).
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
  default: break;
  }
  if (idx < 7 || idx > 19) return false;
  const double *o = kDunOrb;
  for (int r = 7; r < idx; r++) o += 2 * kDunN3[r - 7] + 3 * kDunN6[r - 7];
  if (kDunCen[idx - 7] != 0.0) R.one(1.0 / 3.0, 1.0 / 3.0, kDunCen[idx - 7]);
  for (int k = 0; k < kDunN3[idx - 7]; k++, o += 2) R.orbit3(o[0], o[1]);
  for (int k = 0; k < kDunN6[idx - 7]; k++, o += 3) R.orbit6(o[0], o[1], o[2]);
  return true;
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
int fin_blocks(const std::vector<int> &ntp, std::vector<int> &blk, std::vector<int> &base)
{
  const int np = (int)ntp.size() - 1;
  blk.assign(1, 0);
  base.assign(1, 0);
  int most = 0;
  for (int r = 0; r < np;) {
    const int r0 = r;
    int n = 0;
    while (r < np && r - r0 < kFinRows) {
      const int v = ntp[r + 1] - ntp[r];
      if (r > r0 && n + v > kFinPos) break;
      n += v;
      r++;
    }
    blk.push_back(r);
    base.push_back(base.back() + (n + kFinPos - 1) / kFinPos * kFinPos);
    most = std::max(most, n);
  }
  return most;
}

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
  T.fin_max = fin_blocks(T.node_tri_ptr, T.fin_blk, T.fin_pos);
  T.tri_pos.resize((size_t)3 * ne);
  for (size_t b = 0; b + 1 < T.fin_blk.size(); b++) {
    const int q0 = T.node_tri_ptr[T.fin_blk[b]];
    for (int p = q0; p < T.node_tri_ptr[T.fin_blk[b + 1]]; p++)
      T.tri_pos[(size_t)T.node_tri_k[p] * ne + T.node_tri[p]] = T.fin_pos[b] + (p - q0);
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

void mt_plan(int ne, double tile_bytes, double item_bytes, double limit, int seg,
             MtPlan &p)
{
  const int ng = (ne + 63) / 64;
  p = MtPlan();
  p.ng = ng;
  p.rowoff.assign(ng + 1, 0);
  for (int I = 0; I < ng; I++) p.rowoff[I + 1] = p.rowoff[I] + (ng - I);
  auto nseg = [seg](int n) { return (n + seg - 1) / seg; };
  auto bytes = [&](int I0, int I1) {
    size_t n = 0;
    for (int K = I0; K < I1; K++) n += nseg(ng - K);
    for (int K = I0 + 1; K < ng; K++) n += nseg(std::min(I1, K) - I0);
    return (double)(p.rowoff[I1] - p.rowoff[I0]) * tile_bytes + (double)n * item_bytes;
  };
  p.rows.assign(1, 0);
  p.t0.assign(1, 0);
  p.ioff.assign(1, 0);
  p.koff.assign(1, 0);
  for (int I0 = 0; I0 < ng;) {
    int I1 = std::min(ng, I0 + seg);
    while (I1 < ng && bytes(I0, std::min(ng, I1 + seg)) <= limit) I1 = std::min(ng, I1 + seg);
    const int n0 = (int)(p.item.size() / 4);
    for (int K = I0; K < ng; K++) {
      p.kptr.push_back((int)(p.item.size() / 4) - n0);
      const int c1 = std::min(I1, K);
      for (int M0 = I0; M0 < c1; M0 += seg)
        p.item.insert(p.item.end(), {K, 1, M0, std::min(c1, M0 + seg)});
      if (K < I1)
        for (int M0 = K; M0 < ng; M0 += seg)
          p.item.insert(p.item.end(), {K, 0, M0, std::min(ng, M0 + seg)});
    }
    p.kptr.push_back((int)(p.item.size() / 4) - n0);
    {
      // launch order: the longest items first, so that the short ones fill
      // the last round of waves
      const int ni = (int)(p.item.size() / 4) - n0;
      std::vector<int> ord(ni);
      for (int i = 0; i < ni; i++) ord[i] = i;
      const int *it = p.item.data() + 4 * (size_t)n0;
      std::stable_sort(ord.begin(), ord.end(), [it](int x, int y) {
        return it[4 * x + 3] - it[4 * x + 2] > it[4 * y + 3] - it[4 * y + 2];
      });
      p.order.insert(p.order.end(), ord.begin(), ord.end());
    }
    p.rows.push_back(I1);
    p.t0.push_back(p.rowoff[I1]);
    p.ioff.push_back((int)(p.item.size() / 4));
    p.koff.push_back((int)p.kptr.size());
    p.max_tiles = std::max(p.max_tiles, (size_t)(p.rowoff[I1] - p.rowoff[I0]));
    p.max_items = std::max(p.max_items, p.item.size() / 4 - (size_t)n0);
    I0 = I1;
  }
}

}  // namespace ebh
