/* CPU oracle for the matrix-free layer-potential apply -- TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this file's library (oracle/liboracle.so); the product never does.
 *
 * What pins each part (tests named in tests/test_oracle_*.py):
 *   quadrature tables     equality with the reference tables exported by the
 *                         compiled reference (oracle/_ref) and the golden
 *                         fixture tests/golden/quadrature.npz, plus exactness
 *                         on monomials up to each rule's degree
 *   P0 V, P0xP1 K, M      entrywise against the compiled reference's dense
 *                         getLaplace3d / getLaplace3dMass and the golden
 *                         fixtures (<= 1e-13 relative), plus Gauss's identity
 *                         K*1 = -1/2 M*1 and the unit-sphere eigenvalue 1/5
 *   P0 Kelvin V           entrywise against getElasticity3d and fixtures
 *   Sauter-Schwab maps    polynomial moments against closed forms
 *   P1 blocks (tables)    block sums against the reference reduced formulas,
 *                         full 4-D integration of the maps at high order, and
 *                         symmetry of the assembled P1 single layer
 * No function of this file is left "parity unpinned".
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

#define ORC_PI 3.14159265358979323846
#define MAXQ 80

/* ------------------------------------------------------------------------ */
/* Quadrature                                                               */
/* ------------------------------------------------------------------------ */

/* Sizes of the reference's Gauss-Legendre table (src/quadrature.cpp:286-287). */
static const int kLineSizes[19] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12,
                                   16, 20, 24, 32, 40, 48, 64};

/* Gauss-Legendre on [0,1] by Newton iteration on P_n (long double). */
int orc_gauss01(int n, double *pts, double *wts)
{
  if (n < 1 || n > 128) return -1;
  for (int i = 0; i < (n + 1) / 2; i++) {
    long double x = cosl((long double)ORC_PI * (i + 0.75L) / (n + 0.5L));
    long double dp = 1.0L;
    for (int it = 0; it < 100; it++) {
      long double p0 = 1.0L, p1 = x;
      for (int k = 2; k <= n; k++) {
        long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) { p1 = x; p0 = 1.0L; }
      dp = n * (x * p1 - p0) / (x * x - 1.0L);
      long double dx = p1 / dp;
      x -= dx;
      if (fabsl(dx) < 1e-19L) break;
    }
    {
      long double p0 = 1.0L, p1 = x;
      for (int k = 2; k <= n; k++) {
        long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) { p1 = x; p0 = 1.0L; }
      dp = n * (x * p1 - p0) / (x * x - 1.0L);
    }
    long double w = 2.0L / ((1.0L - x * x) * dp * dp);
    pts[i] = (double)((1.0L - x) / 2.0L);
    pts[n - 1 - i] = (double)((1.0L + x) / 2.0L);
    wts[i] = wts[n - 1 - i] = (double)(w / 2.0L);
  }
  if (n % 2 == 1) pts[n / 2] = 0.5;
  return n;
}

int orc_line_rule(int idx, double *pts, double *wts)
{
  if (idx < 0 || idx >= 19) return -1;
  return orc_gauss01(kLineSizes[idx], pts, wts);
}

/* Published symmetric triangle rules on {x,y >= 0, x+y <= 1}, weights
 * summing to 1/2, indexed like the reference table
 * (src/quadrature.cpp:1047-1048): index d-1 has polynomial degree d.
 * 1, 3, 4 points: Strang & Fix; 6, 12, 13 points: Dunavant (1985);
 * 7 points: Radon's degree-5 rule in closed form; 16..79 points: the
 * Dunavant table below. */
static int push3(double *p, double *w, int n, double a, double wt)
{ /* orbit (a, a), (a, 1-2a), (1-2a, a) */
  double b = 1.0 - 2.0 * a;
  p[2 * n] = a; p[2 * n + 1] = a; w[n++] = wt;
  p[2 * n] = a; p[2 * n + 1] = b; w[n++] = wt;
  p[2 * n] = b; p[2 * n + 1] = a; w[n++] = wt;
  return n;
}
static int push6(double *p, double *w, int n, double a, double b, double wt)
{ /* all permutations of barycentrics (a, b, 1-a-b) */
  double c = 1.0 - a - b;
  double q[6][2] = {{a, b}, {b, c}, {c, a}, {b, a}, {c, b}, {a, c}};
  for (int k = 0; k < 6; k++) {
    p[2 * n] = q[k][0]; p[2 * n + 1] = q[k][1]; w[n++] = wt;
  }
  return n;
}

/* Dunavant (1985) degree 8..20, reference indices 7..19 (16..79 points),
 * weights halved: centroid weight (0: none), then kDunN3 orbits {a, w}
 * (push3) and kDunN6 orbits {a, b, w} (push6), back to back.  The degree-20
 * orbit a = 0.45513668195028 is Dunavant's; the reference's table has
 * 0.45513668695028 there (a transcription error, 5e-9 off the orbit). */
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

int orc_tri_rule(int idx, double *p, double *w)
{
  int n = 0;
  switch (idx) {
  case 0:
    p[0] = p[1] = 1.0 / 3.0; w[0] = 0.5; return 1;
  case 1:
    return push3(p, w, 0, 1.0 / 6.0, 1.0 / 6.0);
  case 2:
    p[0] = p[1] = 1.0 / 3.0; w[0] = -27.0 / 96.0; n = 1;
    return push3(p, w, n, 0.2, 25.0 / 96.0);
  case 3:
    n = push3(p, w, 0, 0.445948490915965, 0.223381589678011 / 2);
    return push3(p, w, n, 0.091576213509771, 0.109951743655322 / 2);
  case 4: {
    double s = sqrt(15.0);
    p[0] = p[1] = 1.0 / 3.0; w[0] = 9.0 / 80.0; n = 1;
    n = push3(p, w, n, (6.0 - s) / 21.0, (155.0 - s) / 2400.0);
    return push3(p, w, n, (6.0 + s) / 21.0, (155.0 + s) / 2400.0);
  }
  case 5:
    n = push3(p, w, 0, 0.249286745170910, 0.116786275726379 / 2);
    n = push3(p, w, n, 0.063089014491502, 0.050844906370207 / 2);
    return push6(p, w, n, 0.053145049844817, 0.310352451033784,
                 0.082851075618374 / 2);
  case 6:
    p[0] = p[1] = 1.0 / 3.0; w[0] = -0.149570044467682 / 2; n = 1;
    n = push3(p, w, n, 0.260345966079040, 0.175615257433208 / 2);
    n = push3(p, w, n, 0.065130102902216, 0.053347235608838 / 2);
    return push6(p, w, n, 0.048690315425316, 0.312865496004874,
                 0.077113760890257 / 2);
  default:
    break;
  }
  if (idx < 7 || idx > 19) return -1;
  {
    const double *o = kDunOrb;
    int r, k;
    for (r = 7; r < idx; r++) o += 2 * kDunN3[r - 7] + 3 * kDunN6[r - 7];
    if (kDunCen[idx - 7] != 0.0) {
      p[0] = p[1] = 1.0 / 3.0;
      w[0] = kDunCen[idx - 7];
      n = 1;
    }
    for (k = 0; k < kDunN3[idx - 7]; k++, o += 2) n = push3(p, w, n, o[0], o[1]);
    for (k = 0; k < kDunN6[idx - 7]; k++, o += 3) n = push6(p, w, n, o[0], o[1], o[2]);
    return n;
  }
}

typedef struct { int n; double p[MAXQ], w[MAXQ]; } line_rule;
typedef struct { int n; double x[MAXQ], y[MAXQ], w[MAXQ]; } tri_rule;

static int load_line(int idx, line_rule *L)
{
  L->n = orc_line_rule(idx, L->p, L->w);
  return L->n > 0 ? 0 : -1;
}
static int load_tri(int idx, tri_rule *T)
{
  double p[2 * MAXQ];
  T->n = orc_tri_rule(idx, p, T->w);
  if (T->n <= 0) return -1;
  for (int i = 0; i < T->n; i++) { T->x[i] = p[2 * i]; T->y[i] = p[2 * i + 1]; }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Small vector helpers                                                     */
/* ------------------------------------------------------------------------ */
static inline double dot3(const double *a, const double *b)
{ return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3(const double *a, const double *b, double *c)
{
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double ncross(const double *a, const double *b)
{ double c[3]; cross3(a, b, c); return sqrt(dot3(c, c)); }
static inline void sub3(const double *a, const double *b, double *c)
{ c[0] = a[0] - b[0]; c[1] = a[1] - b[1]; c[2] = a[2] - b[2]; }
/* d = a*p + b*q + c*r + e*s */
static inline void lin4(double a, const double *p, double b, const double *q,
                        double c, const double *r, double e, const double *s,
                        double *d)
{
  for (int k = 0; k < 3; k++) d[k] = a * p[k] + b * q[k] + c * r[k] + e * s[k];
}

/* ------------------------------------------------------------------------ */
/* Reference integrators restated (FP64)                                    */
/* ------------------------------------------------------------------------ */

/* Coincident single layer, 1-D reduced: src/BEM3dLaplace.cpp:22-52.
 * u = P2-P1, v = P3-P2 of the one triangle. */
static double sl_id(const double *u, const double *v, const line_rule *L)
{
  double c[3]; cross3(u, v, c);
  double jsq = dot3(c, c), s = 0.0, d[3];
  for (int q = 0; q < L->n; q++) {
    double e = L->p[q], t = 0.0;
    for (int k = 0; k < 3; k++) d[k] = e * u[k] + v[k];
    t += 1.0 / sqrt(dot3(d, d));
    for (int k = 0; k < 3; k++) d[k] = e * v[k] + u[k];
    t += 1.0 / sqrt(dot3(d, d));
    for (int k = 0; k < 3; k++) d[k] = e * (u[k] + v[k]) - v[k];
    t += 1.0 / sqrt(dot3(d, d));
    s += L->w[q] * t;
  }
  return jsq * s / ORC_PI / 12.0;
}

/* The five edge-adjacent difference vectors, in the reference's order
 * (src/BEM3dLaplace.cpp:76-99): the first four carry a factor eta2. */
static void edge_dvec(int m, double e2, double e3, const double *u,
                      const double *v, const double *w, double *d)
{
  for (int k = 0; k < 3; k++) {
    double uk = u[k], vk = v[k], wk = w[k];
    switch (m) {
    case 0: d[k] = e2 * (e3 * (uk + wk) - wk) + vk; break;
    case 1: d[k] = -e2 * (uk + vk + e3 * wk) + vk; break;
    case 2: d[k] = e2 * (vk - e3 * (uk + vk)) - wk; break;
    case 3: d[k] = -e2 * (wk + e3 * (uk + vk)) + vk; break;
    default: d[k] = e2 * (uk + wk) + e3 * vk - wk; break;
    }
  }
}

/* Edge-adjacent single layer, 2-D reduced: src/BEM3dLaplace.cpp:55-102. */
static double sl_edge(const double *u, const double *v, const double *w,
                      const line_rule *L)
{
  double J = ncross(u, v) * ncross(u, w), s = 0.0, d[3];
  for (int a = 0; a < L->n; a++)
    for (int b = 0; b < L->n; b++) {
      double e2 = L->p[a], e3 = L->p[b], t = 0.0;
      for (int m = 0; m < 4; m++) {
        edge_dvec(m, e2, e3, u, v, w, d);
        t += 1.0 / sqrt(dot3(d, d));
      }
      t *= e2;
      edge_dvec(4, e2, e3, u, v, w, d);
      t += 1.0 / sqrt(dot3(d, d));
      s += L->w[a] * L->w[b] * t;
    }
  return J * s / ORC_PI / 24.0;
}

/* Vertex-adjacent single layer, 3-D reduced: src/BEM3dLaplace.cpp:105-148. */
static double sl_vertex(const double *u, const double *v, const double *w,
                        const double *z, const line_rule *L)
{
  double J = ncross(u, v) * ncross(w, z), s = 0.0, d[3];
  for (int a = 0; a < L->n; a++)
    for (int b = 0; b < L->n; b++)
      for (int c = 0; c < L->n; c++) {
        double e1 = L->p[a], e2 = L->p[b], e3 = L->p[c], t;
        lin4(1.0, u, e1, v, -e2, w, -e2 * e3, z, d);
        t = 1.0 / sqrt(dot3(d, d));
        lin4(1.0, w, e1, z, -e2, u, -e2 * e3, v, d);
        t += 1.0 / sqrt(dot3(d, d));
        s += L->w[a] * L->w[b] * L->w[c] * e2 * t;
      }
  return J * s / ORC_PI / 12.0;
}

/* Disjoint single layer, tensor triangle rule: src/BEM3dLaplace.cpp:151-192. */
static double sl_far(const double *A, const double *u, const double *v,
                     const double *B, const double *w, const double *z,
                     const tri_rule *T)
{
  double J = ncross(u, v) * ncross(w, z), s = 0.0;
  for (int a = 0; a < T->n; a++) {
    double x[3];
    for (int k = 0; k < 3; k++) x[k] = A[k] + T->x[a] * u[k] + T->y[a] * v[k];
    for (int b = 0; b < T->n; b++) {
      double d[3];
      for (int k = 0; k < 3; k++)
        d[k] = x[k] - B[k] - T->x[b] * w[k] - T->y[b] * z[k];
      s += T->w[a] * T->w[b] / sqrt(dot3(d, d));
    }
  }
  return J * s / ORC_PI / 4.0;
}

/* Edge-adjacent double layer (P0 test x P1 trial): src/BEM3dLaplace.cpp:195-309.
 * The radial variable uses the 2-point Gauss rule and eta1 is the midpoint. */
static void dl_edge(const double *u, const double *v, const double *w,
                    const line_rule *L, const line_rule *L2, double *K)
{
  double c[3], n[3];
  cross3(u, w, c);
  double JT = sqrt(dot3(c, c));
  for (int k = 0; k < 3; k++) n[k] = c[k] / JT;
  double J = ncross(u, v) * JT;
  const double e1 = 0.5;
  double I[3] = {0, 0, 0}, d[3];
  for (int a = 0; a < L->n; a++)
    for (int b = 0; b < L->n; b++) {
      double e2 = L->p[a], e3 = L->p[b];
      double e12 = e1 * e2, e123 = e12 * e3;
      double t[3] = {0, 0, 0};
      for (int m = 0; m < 5; m++) {
        edge_dvec(m, e2, e3, u, v, w, d);
        double r = sqrt(dot3(d, d));
        double H = dot3(d, n) / (r * r * r);
        if (m == 4) { t[0] *= e2; t[1] *= e2; t[2] *= e2; }
        for (int g = 0; g < 2; g++) {
          double P = L2->p[g], W = L2->w[g] * P;
          double f0, f1, f2;
          switch (m) {
          case 0: f0 = 1.0 - P * (1.0 - e123); f1 = P * (1.0 - e12);
                  f2 = P * e12 * (1.0 - e3); break;
          case 1: f0 = 1.0 - P; f1 = P * (1.0 - e123); f2 = P * e123; break;
          case 2: f0 = 1.0 - P; f1 = P * (1.0 - e1); f2 = P * e1; break;
          case 3: f0 = 1.0 - P; f1 = P * (1.0 - e12); f2 = P * e12; break;
          default: f0 = 1.0 - P * (1.0 - e12); f1 = P * (1.0 - e1);
                   f2 = P * e1 * (1.0 - e2); break;
          }
          t[0] += W * f0 * H; t[1] += W * f1 * H; t[2] += W * f2 * H;
        }
      }
      for (int k = 0; k < 3; k++) I[k] += L->w[a] * L->w[b] * t[k];
    }
  for (int k = 0; k < 3; k++) K[k] = J * I[k] / 4.0 / ORC_PI;
}

/* Vertex-adjacent double layer: src/BEM3dLaplace.cpp:312-383. */
static void dl_vertex(const double *u, const double *v, const double *w,
                      const double *z, const line_rule *L,
                      const line_rule *L2, double *K)
{
  double c[3], n[3];
  cross3(w, z, c);
  double JT = sqrt(dot3(c, c));
  for (int k = 0; k < 3; k++) n[k] = c[k] / JT;
  double J = ncross(u, v) * JT;
  double I[3] = {0, 0, 0}, d[3];
  for (int a = 0; a < L->n; a++)
    for (int b = 0; b < L->n; b++)
      for (int cc = 0; cc < L->n; cc++) {
        double e1 = L->p[a], e2 = L->p[b], e3 = L->p[cc], t[3] = {0, 0, 0};
        for (int m = 0; m < 2; m++) {
          if (m == 0) lin4(1.0, u, e1, v, -e2, w, -e2 * e3, z, d);
          else lin4(e2, u, e2 * e3, v, -1.0, w, -e1, z, d);
          double r = sqrt(dot3(d, d));
          double H = dot3(d, n) / (r * r * r);
          for (int g = 0; g < 2; g++) {
            double P = L2->p[g], W = L2->w[g] * P;
            double f0, f1, f2;
            if (m == 0) { f0 = 1.0 - P * e2; f1 = P * e2 * (1.0 - e3);
                          f2 = P * e2 * e3; }
            else { f0 = 1.0 - P; f1 = P * (1.0 - e1); f2 = P * e1; }
            t[0] += W * f0 * H; t[1] += W * f1 * H; t[2] += W * f2 * H;
          }
        }
        double ww = L->w[a] * L->w[b] * L->w[cc] * e2;
        for (int k = 0; k < 3; k++) I[k] += ww * t[k];
      }
  for (int k = 0; k < 3; k++) K[k] = J * I[k] / 4.0 / ORC_PI;
}

/* Disjoint double layer: src/BEM3dLaplace.cpp:386-437. */
static void dl_far(const double *A, const double *u, const double *v,
                   const double *B, const double *w, const double *z,
                   const tri_rule *T, double *K)
{
  double c[3], n[3];
  cross3(w, z, c);
  double JT = sqrt(dot3(c, c));
  for (int k = 0; k < 3; k++) n[k] = c[k] / JT;
  double J = ncross(u, v) * JT, I[3] = {0, 0, 0};
  for (int a = 0; a < T->n; a++) {
    double x[3];
    for (int k = 0; k < 3; k++) x[k] = A[k] + T->x[a] * u[k] + T->y[a] * v[k];
    for (int b = 0; b < T->n; b++) {
      double d[3], e1 = T->x[b], e2 = T->y[b];
      for (int k = 0; k < 3; k++) d[k] = x[k] - B[k] - e1 * w[k] - e2 * z[k];
      double r = sqrt(dot3(d, d));
      double H = T->w[a] * T->w[b] * dot3(d, n) / (r * r * r);
      I[0] += (1.0 - e1 - e2) * H; I[1] += e1 * H; I[2] += e2 * H;
    }
  }
  for (int k = 0; k < 3; k++) K[k] = J * I[k] / ORC_PI / 4.0;
}

/* Kelvin integrands at difference vector d: I, I11, I22, I33, I12, I13, I23
 * (src/BEM3dElasticity.cpp:46-83). */
static inline void kelvin_terms(const double *d, double s, double *I)
{
  double r = sqrt(dot3(d, d)), r3 = r * r * r;
  I[0] += s / r;
  I[1] += s * d[0] * d[0] / r3; I[2] += s * d[1] * d[1] / r3;
  I[3] += s * d[2] * d[2] / r3; I[4] += s * d[0] * d[1] / r3;
  I[5] += s * d[0] * d[2] / r3; I[6] += s * d[1] * d[2] / r3;
}

/* Kelvin reductions: src/BEM3dElasticity.cpp:30-100 (coincident),
 * :103-207 (edge), :210-286 (vertex), :289-348 (disjoint). */
static void ke_id(const double *u, const double *v, const line_rule *L,
                  double *I)
{
  double c[3]; cross3(u, v, c);
  double jsq = dot3(c, c), d[3];
  memset(I, 0, 7 * sizeof(double));
  for (int q = 0; q < L->n; q++) {
    double e = L->p[q];
    for (int k = 0; k < 3; k++) d[k] = e * u[k] + v[k];
    kelvin_terms(d, L->w[q], I);
    for (int k = 0; k < 3; k++) d[k] = e * v[k] + u[k];
    kelvin_terms(d, L->w[q], I);
    for (int k = 0; k < 3; k++) d[k] = e * (u[k] + v[k]) - v[k];
    kelvin_terms(d, L->w[q], I);
  }
  for (int k = 0; k < 7; k++) I[k] *= jsq / ORC_PI / 12.0;
}
static void ke_edge(const double *u, const double *v, const double *w,
                    const line_rule *L, double *I)
{
  double J = ncross(u, v) * ncross(u, w), d[3];
  memset(I, 0, 7 * sizeof(double));
  for (int a = 0; a < L->n; a++)
    for (int b = 0; b < L->n; b++) {
      double e2 = L->p[a], e3 = L->p[b], ww = L->w[a] * L->w[b];
      for (int m = 0; m < 5; m++) {
        edge_dvec(m, e2, e3, u, v, w, d);
        kelvin_terms(d, m < 4 ? ww * e2 : ww, I);
      }
    }
  for (int k = 0; k < 7; k++) I[k] *= J / ORC_PI / 24.0;
}
static void ke_vertex(const double *u, const double *v, const double *w,
                      const double *z, const line_rule *L, double *I)
{
  double J = ncross(u, v) * ncross(w, z), d[3];
  memset(I, 0, 7 * sizeof(double));
  for (int a = 0; a < L->n; a++)
    for (int b = 0; b < L->n; b++)
      for (int c = 0; c < L->n; c++) {
        double e1 = L->p[a], e2 = L->p[b], e3 = L->p[c];
        double ww = L->w[a] * L->w[b] * L->w[c] * e2;
        lin4(1.0, u, e1, v, -e2, w, -e2 * e3, z, d);
        kelvin_terms(d, ww, I);
        lin4(1.0, w, e1, z, -e2, u, -e2 * e3, v, d);
        kelvin_terms(d, ww, I);
      }
  for (int k = 0; k < 7; k++) I[k] *= J / ORC_PI / 12.0;
}
static void ke_far(const double *A, const double *u, const double *v,
                   const double *B, const double *w, const double *z,
                   const tri_rule *T, double *I)
{
  double J = ncross(u, v) * ncross(w, z);
  memset(I, 0, 7 * sizeof(double));
  for (int a = 0; a < T->n; a++) {
    double x[3];
    for (int k = 0; k < 3; k++) x[k] = A[k] + T->x[a] * u[k] + T->y[a] * v[k];
    for (int b = 0; b < T->n; b++) {
      double d[3];
      for (int k = 0; k < 3; k++)
        d[k] = x[k] - B[k] - T->x[b] * w[k] - T->y[b] * z[k];
      kelvin_terms(d, T->w[a] * T->w[b], I);
    }
  }
  for (int k = 0; k < 7; k++) I[k] *= J / ORC_PI / 4.0;
}

/* ------------------------------------------------------------------------ */
/* Pair classification and local geometry (src/BEM3dLaplace.cpp:602-765)   */
/* ------------------------------------------------------------------------ */
typedef struct {
  int cls;            /* 3 coincident, 2 edge, 1 vertex, 0 disjoint */
  int oi[3], oj[3];   /* local vertex orders (global 0-based node ids)    */
  double A[3], B[3], u[3], v[3], w[3], z[3];
  int kscatter[3];    /* reference K scatter targets (0-based nodes)      */
} pair_geom;

static void pair_setup(const double *nodes, const int *elems, int i, int j,
                       pair_geom *g)
{
  const int *fi = elems + 3 * i, *fj = elems + 3 * j;
  double Pi[3][3], Pj[3][3];
  for (int k = 0; k < 3; k++) {
    memcpy(Pi[k], nodes + 3 * (fi[k] - 1), 3 * sizeof(double));
    memcpy(Pj[k], nodes + 3 * (fj[k] - 1), 3 * sizeof(double));
  }
  if (i == j) {
    g->cls = 3;
    for (int k = 0; k < 3; k++) {
      g->oi[k] = g->oj[k] = fi[k] - 1;
      g->u[k] = Pi[1][k] - Pi[0][k];
      g->v[k] = Pi[2][k] - Pi[1][k];
    }
    return;
  }
  int ci[2], cj[2], cnt = 0;
  for (int k = 0; k < 3; k++)
    for (int l = 0; l < 3; l++)
      if (fj[k] == fi[l]) {
        if (cnt < 2) { ci[cnt] = l; cj[cnt] = k; }
        cnt++;
        break;
      }
  if (cnt == 2 && cj[0] == 0 && cj[1] == 2) { /* orient the shared edge */
    int t = cj[0]; cj[0] = cj[1]; cj[1] = t;
    t = ci[0]; ci[0] = ci[1]; ci[1] = t;
  }
  if (cnt == 0) {
    g->cls = 0;
    for (int k = 0; k < 3; k++) {
      g->A[k] = Pi[0][k]; g->B[k] = Pj[0][k];
      g->u[k] = Pi[1][k] - Pi[0][k]; g->v[k] = Pi[2][k] - Pi[0][k];
      g->w[k] = Pj[1][k] - Pj[0][k]; g->z[k] = Pj[2][k] - Pj[0][k];
      g->oi[k] = fi[k] - 1; g->oj[k] = fj[k] - 1;
      g->kscatter[k] = fj[k] - 1;
    }
  } else if (cnt == 1) {
    int c = ci[0], d = cj[0];
    g->cls = 1;
    for (int k = 0; k < 3; k++) {
      g->u[k] = Pi[(c + 1) % 3][k] - Pi[c][k];
      g->v[k] = Pi[(c + 2) % 3][k] - Pi[(c + 1) % 3][k];
      g->w[k] = Pj[(d + 1) % 3][k] - Pj[d][k];
      g->z[k] = Pj[(d + 2) % 3][k] - Pj[(d + 1) % 3][k];
    }
    for (int k = 0; k < 3; k++) {
      g->oi[k] = fi[(c + k) % 3] - 1;
      g->oj[k] = fj[(d + k) % 3] - 1;
      g->kscatter[k] = g->oj[k];
    }
  } else {
    int remi = 3 - ci[0] - ci[1], remj = 3 - cj[0] - cj[1];
    g->cls = 2;
    for (int k = 0; k < 3; k++) {
      g->u[k] = Pj[cj[1]][k] - Pj[cj[0]][k];
      g->v[k] = Pi[remi][k] - Pj[cj[1]][k];
      g->w[k] = Pj[remj][k] - Pj[cj[1]][k];
    }
    g->oi[0] = fi[ci[0]] - 1; g->oi[1] = fi[ci[1]] - 1; g->oi[2] = fi[remi] - 1;
    g->oj[0] = fj[cj[0]] - 1; g->oj[1] = fj[cj[1]] - 1; g->oj[2] = fj[remj] - 1;
    for (int k = 0; k < 3; k++) g->kscatter[k] = g->oj[k];
  }
}

/* ------------------------------------------------------------------------ */
/* Sauter-Schwab maps (reference triangle T = {0 <= s2 <= s1 <= 1})        */
/* ------------------------------------------------------------------------ */
/* Vertex-adjacent (2 regions) and edge-adjacent (5 regions): the maps of
 * Sauter & Schwab, "Boundary Element Methods", sec. 5.2.1, written so that
 * (x - y)/h reproduces the reference difference vectors above.
 * Coincident: z = s - t runs over the hexagon T - T along six rays
 * z = rho*zhat_m(eta); for fixed z the inner point s ranges over the triangle
 * T n (T + z), whose area is (1-rho)^2/2, parameterised by barycentric
 * (l1, l2) on the unit simplex.
 * ss_map returns s, t, the Jacobian and the homogeneity factor h with
 * x - y = h * d_m(outer). */
static void ss_map(int cls, int m, const double *in, const double *out,
                   double *s, double *t, double *jac, double *h)
{
  if (cls == 1) { /* vertex: inner xi; outer e1, e2, e3 */
    double xi = in[0], e1 = out[0], e2 = out[1], e3 = out[2];
    double a0 = xi, a1 = xi * e1, b0 = xi * e2, b1 = xi * e2 * e3;
    if (m == 0) { s[0] = a0; s[1] = a1; t[0] = b0; t[1] = b1; }
    else { s[0] = b0; s[1] = b1; t[0] = a0; t[1] = a1; }
    *jac = xi * xi * xi * e2;
    *h = xi;
  } else if (cls == 2) { /* edge: inner xi, e1; outer e2, e3 */
    double xi = in[0], e1 = in[1], e2 = out[0], e3 = out[1];
    switch (m) {
    case 0: /* difference vector 0: eta2(eta3(u+w)-w)+v */
      s[0] = xi; s[1] = xi * e1;
      t[0] = xi * (1 - e1 * e2 * e3); t[1] = xi * e1 * e2 * (1 - e3); break;
    case 1: /* -eta2(u+v+eta3 w)+v */
      s[0] = xi * (1 - e1 * e2); s[1] = xi * e1 * (1 - e2);
      t[0] = xi; t[1] = xi * e1 * e2 * e3; break;
    case 2: /* eta2(v-eta3(u+v))-w */
      s[0] = xi * (1 - e1 * e2 * e3); s[1] = xi * e1 * e2 * (1 - e3);
      t[0] = xi; t[1] = xi * e1; break;
    case 3: /* -eta2(w+eta3(u+v))+v */
      s[0] = xi * (1 - e1 * e2 * e3); s[1] = xi * e1 * (1 - e2 * e3);
      t[0] = xi; t[1] = xi * e1 * e2; break;
    default: /* eta2(u+w)+eta3 v-w, no eta2 factor */
      s[0] = xi; s[1] = xi * e1 * e3;
      t[0] = xi * (1 - e1 * e2); t[1] = xi * e1 * (1 - e2); break;
    }
    *jac = xi * xi * xi * e1 * e1 * (m < 4 ? e2 : 1.0);
    *h = xi * e1;
  } else { /* coincident: inner rho, l1, l2; outer eta */
    double rho = in[0], l1 = in[1], l2 = in[2], eta = out[0];
    double zh[2];
    switch (m % 3) {
    case 0: zh[0] = eta; zh[1] = 1.0; break;
    case 1: zh[0] = 1.0; zh[1] = eta; break;
    default: zh[0] = eta; zh[1] = eta - 1.0; break;
    }
    if (m >= 3) { zh[0] = -zh[0]; zh[1] = -zh[1]; }
    double z1 = rho * zh[0], z2 = rho * zh[1];
    double a = z2 > 0 ? z2 : 0.0;
    double b = 1.0 + (z1 < 0 ? z1 : 0.0);
    double c = (z1 - z2) > 0 ? (z1 - z2) : 0.0;
    double V1[2] = {a + c, a}, V2[2] = {b, a}, V3[2] = {b, b - c};
    s[0] = V1[0] + l1 * (V2[0] - V1[0]) + l2 * (V3[0] - V1[0]);
    s[1] = V1[1] + l1 * (V2[1] - V1[1]) + l2 * (V3[1] - V1[1]);
    t[0] = s[0] - z1;
    t[1] = s[1] - z2;
    double side = b - a - c;
    *jac = rho * side * side;
    *h = rho;
  }
}

/* P1 shape functions on the Sauter triangle, vertex order (a0, a1, a2). */
static inline void sauter_phi(const double *s, double *f)
{ f[0] = 1.0 - s[0]; f[1] = s[0] - s[1]; f[2] = s[1]; }

static int ss_nregions(int cls) { return cls == 1 ? 2 : cls == 2 ? 5 : 3; }
static int ss_nouter(int cls, int nq)
{ return cls == 1 ? nq * nq * nq : cls == 2 ? nq * nq : nq; }

/* Reduced weight tables: for each region m and outer point q,
 * W[(m*nout + q)*9 + 3k + l] = sum over the inner variables of
 * jac / h^deg * phi_k(s) phi_l(t), integrated exactly (inner polynomial
 * degree <= 5 is integrated by 4-point Gauss and a degree-4 simplex rule).
 * The coincident table folds the opposite rays m and m+3 together since
 * every kernel used is even in x - y there (the double layer vanishes). */
int orc_ss_table(int cls, int kdeg, int nq, double *W)
{
  line_rule L, G;
  double gp[8], gw[8];
  if (orc_gauss01(nq, L.p, L.w) < 0) return -1;
  L.n = nq;
  orc_gauss01(4, gp, gw);
  memcpy(G.p, gp, sizeof(gp)); memcpy(G.w, gw, sizeof(gw)); G.n = 4;
  tri_rule S;
  load_tri(3, &S);
  int nreg = ss_nregions(cls), nout = ss_nouter(cls, nq);
  memset(W, 0, sizeof(double) * 9 * nreg * nout);
  for (int q = 0; q < nout; q++) {
    double out[3];
    if (cls == 1) {
      out[0] = L.p[q / (nq * nq)]; out[1] = L.p[(q / nq) % nq]; out[2] = L.p[q % nq];
    } else if (cls == 2) {
      out[0] = L.p[q / nq]; out[1] = L.p[q % nq];
    } else out[0] = L.p[q];
    int mreg = cls == 3 || cls == 0 ? 6 : nreg;
    for (int m = 0; m < mreg; m++) {
      double *Wm = W + ((m % nreg) * nout + q) * 9;
      if (cls == 1) {
        for (int g = 0; g < G.n; g++) {
          double in[1] = {G.p[g]}, s[2], t[2], jac, h, fs[3], ft[3];
          ss_map(cls, m, in, out, s, t, &jac, &h);
          sauter_phi(s, fs); sauter_phi(t, ft);
          double c = G.w[g] * jac / pow(h, kdeg);
          for (int k = 0; k < 3; k++)
            for (int l = 0; l < 3; l++) Wm[3 * k + l] += c * fs[k] * ft[l];
        }
      } else if (cls == 2) {
        for (int g = 0; g < G.n; g++)
          for (int g2 = 0; g2 < G.n; g2++) {
            double in[2] = {G.p[g], G.p[g2]}, s[2], t[2], jac, h, fs[3], ft[3];
            ss_map(cls, m, in, out, s, t, &jac, &h);
            sauter_phi(s, fs); sauter_phi(t, ft);
            double c = G.w[g] * G.w[g2] * jac / pow(h, kdeg);
            for (int k = 0; k < 3; k++)
              for (int l = 0; l < 3; l++) Wm[3 * k + l] += c * fs[k] * ft[l];
          }
      } else {
        for (int g = 0; g < G.n; g++)
          for (int r = 0; r < S.n; r++) {
            double in[3] = {G.p[g], S.x[r], S.y[r]}, s[2], t[2], jac, h;
            double fs[3], ft[3];
            ss_map(3, m, in, out, s, t, &jac, &h);
            sauter_phi(s, fs); sauter_phi(t, ft);
            double c = G.w[g] * S.w[r] * jac / pow(h, kdeg);
            for (int k = 0; k < 3; k++)
              for (int l = 0; l < 3; l++) Wm[3 * k + l] += c * fs[k] * ft[l];
          }
      }
    }
  }
  return nreg * nout;
}

/* Difference vector d_m at outer point q (same formulas as the reference
 * reductions; coincident and vertex-adjacent region 1 keep the x - y sign). */
static void ss_dvec(int cls, int m, const double *out, const pair_geom *g,
                    double *d)
{
  if (cls == 1) {
    double e1 = out[0], e2 = out[1], e3 = out[2];
    if (m == 0) lin4(1.0, g->u, e1, g->v, -e2, g->w, -e2 * e3, g->z, d);
    else lin4(e2, g->u, e2 * e3, g->v, -1.0, g->w, -e1, g->z, d);
  } else if (cls == 2) {
    edge_dvec(m, out[0], out[1], g->u, g->v, g->w, d);
  } else {
    double e = out[0];
    for (int k = 0; k < 3; k++) {
      double uk = g->u[k], vk = g->v[k];
      d[k] = m == 0 ? e * uk + vk : m == 1 ? uk + e * vk : e * (uk + vk) - vk;
    }
  }
}

/* Table-driven near-field pair block.  op 0: SL (9), op 1: DL (9),
 * op 2: Kelvin (63). */
static void near_block_tables(int op, const pair_geom *g, int nq,
                              const double *W, const double *WD, double *out)
{
  int cls = g->cls == 3 ? 3 : g->cls == 2 ? 2 : 1;
  int nreg = ss_nregions(cls), nout = ss_nouter(cls, nq);
  double gp[MAXQ], gw[MAXQ];
  orc_gauss01(nq, gp, gw);
  double J, n[3] = {0, 0, 0}, c[3];
  if (cls == 3) { J = ncross(g->u, g->v); J = J * J; }
  else if (cls == 2) {
    cross3(g->u, g->w, c);
    double jt = sqrt(dot3(c, c));
    for (int k = 0; k < 3; k++) n[k] = c[k] / jt;
    J = ncross(g->u, g->v) * jt;
  } else {
    cross3(g->w, g->z, c);
    double jt = sqrt(dot3(c, c));
    for (int k = 0; k < 3; k++) n[k] = c[k] / jt;
    J = ncross(g->u, g->v) * jt;
  }
  int nval = op == 2 ? 63 : 9;
  memset(out, 0, sizeof(double) * nval);
  if (op == 1 && cls == 3) return; /* (x-y).n_y = 0 on a flat panel */
  for (int q = 0; q < nout; q++) {
    double o[3], wq;
    if (cls == 1) {
      int a = q / (nq * nq), b = (q / nq) % nq, cc = q % nq;
      o[0] = gp[a]; o[1] = gp[b]; o[2] = gp[cc]; wq = gw[a] * gw[b] * gw[cc];
    } else if (cls == 2) {
      int a = q / nq, b = q % nq;
      o[0] = gp[a]; o[1] = gp[b]; wq = gw[a] * gw[b];
    } else { o[0] = gp[q]; wq = gw[q]; }
    for (int m = 0; m < nreg; m++) {
      double d[3];
      ss_dvec(cls, m, o, g, d);
      double r = sqrt(dot3(d, d));
      if (op == 0) {
        const double *Wm = W + (m * nout + q) * 9;
        double kv = wq / r;
        for (int e = 0; e < 9; e++) out[e] += Wm[e] * kv;
      } else if (op == 1) {
        const double *Wm = WD + (m * nout + q) * 9;
        double kv = wq * dot3(d, n) / (r * r * r);
        for (int e = 0; e < 9; e++) out[e] += Wm[e] * kv;
      } else {
        const double *Wm = W + (m * nout + q) * 9;
        double kv[7] = {0, 0, 0, 0, 0, 0, 0};
        kelvin_terms(d, wq, kv);
        for (int e = 0; e < 9; e++)
          for (int cc = 0; cc < 7; cc++) out[7 * e + cc] += Wm[e] * kv[cc];
      }
    }
  }
  for (int e = 0; e < nval; e++) out[e] *= J / (4.0 * ORC_PI);
}

/* Far-field pair block with P1 shape functions (1 - x1 - x2, x1, x2). */
static void far_block(int op, const pair_geom *g, const tri_rule *T,
                      double *out)
{
  double c[3], n[3];
  cross3(g->w, g->z, c);
  double jt = sqrt(dot3(c, c));
  for (int k = 0; k < 3; k++) n[k] = c[k] / jt;
  double J = ncross(g->u, g->v) * jt;
  int nval = op == 2 ? 63 : 9;
  memset(out, 0, sizeof(double) * nval);
  for (int a = 0; a < T->n; a++) {
    double x[3], fa[3] = {1.0 - T->x[a] - T->y[a], T->x[a], T->y[a]};
    for (int k = 0; k < 3; k++)
      x[k] = g->A[k] + T->x[a] * g->u[k] + T->y[a] * g->v[k];
    for (int b = 0; b < T->n; b++) {
      double d[3], e1 = T->x[b], e2 = T->y[b];
      double fb[3] = {1.0 - e1 - e2, e1, e2};
      for (int k = 0; k < 3; k++)
        d[k] = x[k] - g->B[k] - e1 * g->w[k] - e2 * g->z[k];
      double r = sqrt(dot3(d, d)), ww = T->w[a] * T->w[b];
      if (op == 0) {
        double kv = ww / r;
        for (int k = 0; k < 3; k++)
          for (int l = 0; l < 3; l++) out[3 * k + l] += fa[k] * fb[l] * kv;
      } else if (op == 1) {
        double kv = ww * dot3(d, n) / (r * r * r);
        for (int k = 0; k < 3; k++)
          for (int l = 0; l < 3; l++) out[3 * k + l] += fa[k] * fb[l] * kv;
      } else {
        double kv[7] = {0, 0, 0, 0, 0, 0, 0};
        kelvin_terms(d, ww, kv);
        for (int k = 0; k < 3; k++)
          for (int l = 0; l < 3; l++)
            for (int cc = 0; cc < 7; cc++)
              out[7 * (3 * k + l) + cc] += fa[k] * fb[l] * kv[cc];
      }
    }
  }
  for (int e = 0; e < nval; e++) out[e] *= J / (4.0 * ORC_PI);
}

/* ------------------------------------------------------------------------ */
/* Shared context for assembly and applies                                  */
/* ------------------------------------------------------------------------ */
typedef struct {
  line_rule L, L2;
  tri_rule T;
  int nq;
  double *W[4], *WD[4]; /* indexed by cls 1..3 */
} ctx_t;

static int ctx_init(ctx_t *c, int line_idx, int tri_idx, int tables)
{
  memset(c, 0, sizeof(*c));
  if (load_line(line_idx, &c->L) || load_line(1, &c->L2) ||
      load_tri(tri_idx, &c->T))
    return -1;
  c->nq = c->L.n;
  if (tables) {
    for (int cls = 1; cls <= 3; cls++) {
      size_t sz = 9 * (size_t)ss_nregions(cls) * ss_nouter(cls, c->nq);
      c->W[cls] = (double *)malloc(sz * sizeof(double));
      c->WD[cls] = (double *)malloc(sz * sizeof(double));
      orc_ss_table(cls, 1, c->nq, c->W[cls]);
      orc_ss_table(cls, 2, c->nq, c->WD[cls]);
    }
  }
  return 0;
}
static void ctx_free(ctx_t *c)
{
  for (int k = 0; k < 4; k++) { free(c->W[k]); free(c->WD[k]); }
}

/* Block of a pair in pair-local orders via the tables (near) or the tensor
 * rule (far). */
static void pair_block(const ctx_t *c, int op, const pair_geom *g, double *out)
{
  if (g->cls == 0) far_block(op, g, &c->T, out);
  else {
    int cls = g->cls == 3 ? 3 : g->cls == 2 ? 2 : 1;
    near_block_tables(op, g, c->nq, c->W[cls], c->WD[cls], out);
  }
}

/* ------------------------------------------------------------------------ */
/* Surface curls, src/BEM3dLaplace.cpp:489-558: the 2x2 metric system of the  */
/* edges e1 = P2 - P1, e2 = P3 - P1 against b = e x n per edge (:523-547).    */
/* ------------------------------------------------------------------------ */
int orc_surface_curls(int np, const double *nodes, int ne, const int *elems,
                      double *curl)
{
  (void)np;
  for (int i = 0; i < ne; i++) {
    const double *P[3];
    for (int k = 0; k < 3; k++) P[k] = nodes + 3 * (elems[3 * i + k] - 1);
    double e1[3], e2[3], n[3];
    for (int k = 0; k < 3; k++) {
      e1[k] = P[1][k] - P[0][k];
      e2[k] = P[2][k] - P[0][k];
    }
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    const double J = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    double a11 = 0, a12 = 0, a22 = 0;
    for (int k = 0; k < 3; k++) {
      n[k] /= J;
      a11 += e1[k] * e1[k];
      a22 += e2[k] * e2[k];
      a12 += e1[k] * e2[k];
    }
    const double det = a11 * a22 - a12 * a12;
    /* rows c = x, y, z of b1 = e1 x n and b2 = e2 x n */
    const double b1[3] = {-e1[1] * n[2] + e1[2] * n[1], e1[0] * n[2] - e1[2] * n[0],
                          -e1[0] * n[1] + e1[1] * n[0]};
    const double b2[3] = {-e2[1] * n[2] + e2[2] * n[1], e2[0] * n[2] - e2[2] * n[0],
                          -e2[0] * n[1] + e2[1] * n[0]};
    for (int c = 0; c < 3; c++) {
      const double y1 = (b1[c] * a22 - b2[c] * a12) / det;
      const double y2 = (a11 * b2[c] - a12 * b1[c]) / det;
      double *o = curl + (3 * (size_t)i + c) * 3;
      o[0] = -y1 - y2;
      o[1] = y1;
      o[2] = y2;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Representation formula, src/BEM3dLaplace.cpp:784-843: per triangle j and */
/* point, sum_k w_k [t_j/r - u_h(y_k) (x - y_k).n_j/r^3], times J_j/(4 pi).  */
/* ------------------------------------------------------------------------ */
int orc_evaluate(int np, const double *nodes, int ne, const int *elems,
                 int npts, const double *points, const double *t,
                 const double *u, int tri_idx, double *result, int nthreads)
{
  (void)np;
  double qp[2 * MAXQ], qw[MAXQ];
  const int nq = orc_tri_rule(tri_idx, qp, qw);
  if (nq <= 0) return -1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < npts; i++) {
    const double *x = points + 3 * (size_t)i;
    double acc = 0.0;
    for (int j = 0; j < ne; j++) {
      const int a = elems[3 * j] - 1, b = elems[3 * j + 1] - 1, c = elems[3 * j + 2] - 1;
      double A[3], U[3], V[3], n[3];
      for (int d = 0; d < 3; d++) {
        A[d] = nodes[3 * a + d];
        U[d] = nodes[3 * b + d] - A[d];
        V[d] = nodes[3 * c + d] - A[d];
      }
      n[0] = U[1] * V[2] - U[2] * V[1];
      n[1] = U[2] * V[0] - U[0] * V[2];
      n[2] = U[0] * V[1] - U[1] * V[0];
      const double J = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      for (int d = 0; d < 3; d++) n[d] /= J;
      double I = 0.0;
      for (int k = 0; k < nq; k++) {
        const double s1 = qp[2 * k], s2 = qp[2 * k + 1];
        double r[3];
        for (int d = 0; d < 3; d++) r[d] = x[d] - A[d] - U[d] * s1 - V[d] * s2;
        const double rn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        const double uh = u[a] + (u[b] - u[a]) * s1 + (u[c] - u[a]) * s2;
        I += qw[k] * t[j] / rn;
        I -= qw[k] * uh * (r[0] * n[0] + r[1] * n[1] + r[2] * n[2]) / (rn * rn * rn);
      }
      acc += J * I / 4.0 / ORC_PI;
    }
    result[i] = acc;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Dense P0 assembly restating getLaplace3d / getElasticity3d                */
/* ------------------------------------------------------------------------ */
int orc_laplace_dense(int np, const double *nodes, int ne, const int *elems,
                      int line_idx, int tri_idx, double *V, double *K,
                      double *M)
{
  ctx_t c;
  if (ctx_init(&c, line_idx, tri_idx, 0)) return -1;
  if (K) memset(K, 0, sizeof(double) * (size_t)ne * np);
  if (M) memset(M, 0, sizeof(double) * (size_t)ne * np);
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < ne; i++) {
    const int *fi = elems + 3 * i;
    if (M) {
      double a[3], b[3];
      sub3(nodes + 3 * (fi[1] - 1), nodes + 3 * (fi[0] - 1), a);
      sub3(nodes + 3 * (fi[2] - 1), nodes + 3 * (fi[0] - 1), b);
      double Ji = ncross(a, b);
      for (int k = 0; k < 3; k++) M[(size_t)i * np + fi[k] - 1] += Ji / 6.0;
    }
    for (int j = 0; j < ne; j++) {
      pair_geom g;
      pair_setup(nodes, elems, i, j, &g);
      double v = 0.0, k3[3] = {0, 0, 0};
      switch (g.cls) {
      case 3: v = sl_id(g.u, g.v, &c.L); break;
      case 0:
        v = sl_far(g.A, g.u, g.v, g.B, g.w, g.z, &c.T);
        if (K) dl_far(g.A, g.u, g.v, g.B, g.w, g.z, &c.T, k3);
        break;
      case 1:
        v = sl_vertex(g.u, g.v, g.w, g.z, &c.L);
        if (K) dl_vertex(g.u, g.v, g.w, g.z, &c.L, &c.L2, k3);
        break;
      default:
        v = sl_edge(g.u, g.v, g.w, &c.L);
        if (K) dl_edge(g.u, g.v, g.w, &c.L, &c.L2, k3);
        break;
      }
      if (V) V[(size_t)i * ne + j] = v;
      if (K && g.cls != 3)
        for (int k = 0; k < 3; k++) K[(size_t)i * np + g.kscatter[k]] += k3[k];
    }
  }
  ctx_free(&c);
  return 0;
}

static void kelvin_p0(const ctx_t *c, const pair_geom *g, double *I)
{
  switch (g->cls) {
  case 3: ke_id(g->u, g->v, &c->L, I); break;
  case 0: ke_far(g->A, g->u, g->v, g->B, g->w, g->z, &c->T, I); break;
  case 1: ke_vertex(g->u, g->v, g->w, g->z, &c->L, I); break;
  default: ke_edge(g->u, g->v, g->w, &c->L, I); break;
  }
}

/* Kelvin block from the 7 integrals (src/BEM3dElasticity.cpp:612,838-846). */
static void kelvin_block(const double *I, double nu, double *B /* 3x3 */)
{
  double f = (1.0 + nu) / 2.0 / (1.0 - nu), t = 3.0 - 4.0 * nu;
  B[0] = f * (t * I[0] + I[1]); B[4] = f * (t * I[0] + I[2]);
  B[8] = f * (t * I[0] + I[3]);
  B[1] = B[3] = f * I[4]; B[2] = B[6] = f * I[5]; B[5] = B[7] = f * I[6];
}

int orc_kelvin_dense(int np, const double *nodes, int ne, const int *elems,
                     double nu, int line_idx, int tri_idx, double *V)
{
  (void)np;
  ctx_t c;
  if (ctx_init(&c, line_idx, tri_idx, 0)) return -1;
  size_t n3 = 3 * (size_t)ne;
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < ne; i++)
    for (int j = 0; j < ne; j++) {
      pair_geom g;
      double I[7], B[9];
      pair_setup(nodes, elems, i, j, &g);
      kelvin_p0(&c, &g, I);
      kelvin_block(I, nu, B);
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
          V[(a * (size_t)ne + i) * n3 + b * (size_t)ne + j] = B[3 * a + b];
    }
  ctx_free(&c);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Row-sampled matrix-free applies                                          */
/* ------------------------------------------------------------------------ */
static int local_index(const int *fi, int node0)
{
  for (int k = 0; k < 3; k++) if (fi[k] - 1 == node0) return k;
  return -1;
}

/* Contribution of test triangle i against every trial triangle: out[3*nc]
 * indexed by i's own local vertex k (P1 test) or out[c] (P0 test, k=0),
 * nc = 1 (Laplace) or 3 (Kelvin components). */
static void tri_row(const ctx_t *c, int op, int test_basis, int trial_basis,
                    int np, const double *nodes, int ne, const int *elems,
                    double nu, const double *x, int i, double *out)
{
  const int *fi = elems + 3 * i;
  int nc = op == 2 ? 3 : 1;
  int ntrial = trial_basis ? np : ne;
  memset(out, 0, sizeof(double) * 9);
  for (int j = 0; j < ne; j++) {
    pair_geom g;
    pair_setup(nodes, elems, i, j, &g);
    const int *fj = elems + 3 * j;
    if (op == 0 && test_basis == 0 && trial_basis == 0) {
      /* reference P0 single layer */
      double v;
      switch (g.cls) {
      case 3: v = sl_id(g.u, g.v, &c->L); break;
      case 0: v = sl_far(g.A, g.u, g.v, g.B, g.w, g.z, &c->T); break;
      case 1: v = sl_vertex(g.u, g.v, g.w, g.z, &c->L); break;
      default: v = sl_edge(g.u, g.v, g.w, &c->L); break;
      }
      out[0] += v * x[j];
      continue;
    }
    if (op == 1 && test_basis == 0) {
      /* reference P0-test x P1-trial double layer */
      double k3[3] = {0, 0, 0};
      switch (g.cls) {
      case 3: break;
      case 0: dl_far(g.A, g.u, g.v, g.B, g.w, g.z, &c->T, k3); break;
      case 1: dl_vertex(g.u, g.v, g.w, g.z, &c->L, &c->L2, k3); break;
      default: dl_edge(g.u, g.v, g.w, &c->L, &c->L2, k3); break;
      }
      if (g.cls != 3)
        for (int k = 0; k < 3; k++) out[0] += k3[k] * x[g.kscatter[k]];
      continue;
    }
    if (op == 2 && test_basis == 0 && trial_basis == 0) {
      /* reference P0 Kelvin */
      double I[7], B[9];
      kelvin_p0(c, &g, I);
      kelvin_block(I, nu, B);
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
          out[a] += B[3 * a + b] * x[(size_t)b * ne + j];
      continue;
    }
    /* P1 test and/or trial via the pair block */
    double blk[63];
    pair_block(c, op, &g, blk);
    for (int kk = 0; kk < 3; kk++) {
      int ki = test_basis ? local_index(fi, g.oi[kk]) : 0;
      for (int ll = 0; ll < 3; ll++) {
        int col = trial_basis ? g.oj[ll] : j;
        if (op == 2) {
          double I[7], B[9];
          for (int cc = 0; cc < 7; cc++) I[cc] = blk[7 * (3 * kk + ll) + cc];
          kelvin_block(I, nu, B);
          for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
              out[3 * ki + a] += B[3 * a + b] * x[(size_t)b * ntrial + col];
        } else {
          out[3 * ki] += blk[3 * kk + ll] * x[col];
        }
      }
    }
    (void)fj;
  }
}

int orc_apply(int op, int test_basis, int trial_basis, int np,
              const double *nodes, int ne, const int *elems, int line_idx,
              int tri_idx, double nu, const double *x, int nrows,
              const int *rows, double *y, int nthreads)
{
  if (op == 1 && trial_basis != 1) return -2;
  if (op == 2 && test_basis != trial_basis) return -2;
  ctx_t c;
  if (ctx_init(&c, line_idx, tri_idx, 1)) return -1;
  int nc = op == 2 ? 3 : 1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  if (!test_basis) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int r = 0; r < nrows; r++) {
      double out[9];
      tri_row(&c, op, test_basis, trial_basis, np, nodes, ne, elems, nu, x,
              rows[r], out);
      for (int a = 0; a < nc; a++) y[(size_t)a * nrows + r] = out[a];
    }
  } else {
    /* node rows: sum the node's share over the triangles containing it */
#pragma omp parallel for schedule(dynamic, 1)
    for (int r = 0; r < nrows; r++) {
      int n0 = rows[r];
      double acc[3] = {0, 0, 0};
      for (int i = 0; i < ne; i++) {
        int k = local_index(elems + 3 * i, n0);
        if (k < 0) continue;
        double out[9];
        tri_row(&c, op, test_basis, trial_basis, np, nodes, ne, elems, nu, x,
                i, out);
        for (int a = 0; a < nc; a++) acc[a] += out[3 * k + a];
      }
      for (int a = 0; a < nc; a++) y[(size_t)a * nrows + r] = acc[a];
    }
  }
  ctx_free(&c);
  return 0;
}

int orc_pair_block(int op, int np, const double *nodes, int ne,
                   const int *elems, int i, int j, int line_idx, int tri_idx,
                   int use_tables, double *out, int *order_i, int *order_j)
{
  (void)np; (void)ne;
  ctx_t c;
  if (ctx_init(&c, line_idx, tri_idx, use_tables)) return -1;
  pair_geom g;
  pair_setup(nodes, elems, i, j, &g);
  for (int k = 0; k < 3; k++) { order_i[k] = g.oi[k]; order_j[k] = g.oj[k]; }
  if (use_tables) pair_block(&c, op, &g, out);
  else {
    memset(out, 0, sizeof(double) * (op == 2 ? 63 : 9));
    if (op == 0) {
      switch (g.cls) {
      case 3: out[0] = sl_id(g.u, g.v, &c.L); break;
      case 0: out[0] = sl_far(g.A, g.u, g.v, g.B, g.w, g.z, &c.T); break;
      case 1: out[0] = sl_vertex(g.u, g.v, g.w, g.z, &c.L); break;
      default: out[0] = sl_edge(g.u, g.v, g.w, &c.L); break;
      }
    } else if (op == 1) {
      switch (g.cls) {
      case 3: break;
      case 0: dl_far(g.A, g.u, g.v, g.B, g.w, g.z, &c.T, out); break;
      case 1: dl_vertex(g.u, g.v, g.w, g.z, &c.L, &c.L2, out); break;
      default: dl_edge(g.u, g.v, g.w, &c.L, &c.L2, out); break;
      }
    } else kelvin_p0(&c, &g, out);
  }
  ctx_free(&c);
  return g.cls;
}

/* ------------------------------------------------------------------------ */
/* Full 4-D integration of the maps (test only)                             */
/* ------------------------------------------------------------------------ */
double orc_ss_full(int cls, int kind, const int *e, const double *geom, int n)
{
  double gp[128], gw[128];
  if (orc_gauss01(n, gp, gw) < 0) return NAN;
  int nreg = cls == 0 ? 6 : cls == 1 ? 5 : 2;
  int mcls = cls == 0 ? 3 : cls == 1 ? 2 : 1;
  const double *u = geom, *v = geom + 3, *w = geom + 6, *z = geom + 9;
  double total = 0.0;
  for (int m = 0; m < nreg; m++)
    for (int a = 0; a < n; a++)
      for (int b = 0; b < n; b++)
        for (int c = 0; c < n; c++)
          for (int d = 0; d < n; d++) {
            double in[3], out[3], wt = gw[a] * gw[b] * gw[c] * gw[d], s[2],
                   t[2], jac, h;
            if (mcls == 1) {
              in[0] = gp[a]; out[0] = gp[b]; out[1] = gp[c]; out[2] = gp[d];
            } else if (mcls == 2) {
              in[0] = gp[a]; in[1] = gp[b]; out[0] = gp[c]; out[1] = gp[d];
            } else { /* conical product on the simplex for (l1, l2) */
              in[0] = gp[a]; in[1] = gp[b]; in[2] = gp[c] * (1.0 - gp[b]);
              wt *= (1.0 - gp[b]);
              out[0] = gp[d];
            }
            ss_map(mcls, m, in, out, s, t, &jac, &h);
            double f;
            if (kind == 0) {
              f = pow(s[0], e[0]) * pow(s[1], e[1]) * pow(t[0], e[2]) *
                  pow(t[1], e[3]);
            } else {
              double dx[3], fs[3], ft[3];
              for (int k = 0; k < 3; k++) {
                if (mcls == 1)
                  dx[k] = s[0] * u[k] + s[1] * v[k] - t[0] * w[k] - t[1] * z[k];
                else if (mcls == 2)
                  dx[k] = (s[0] - t[0]) * u[k] + s[1] * v[k] - t[1] * w[k];
                else
                  dx[k] = (s[0] - t[0]) * u[k] + (s[1] - t[1]) * v[k];
              }
              sauter_phi(s, fs); sauter_phi(t, ft);
              f = fs[e[0]] * ft[e[1]] / sqrt(dot3(dx, dx));
            }
            total += wt * jac * f;
          }
  return total;
}
