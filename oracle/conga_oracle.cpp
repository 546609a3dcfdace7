// conga_oracle.cpp -- CPU oracle for the SuperConga self-consistency path.
//
// TEST INFRASTRUCTURE ONLY (see conga_oracle.h).  A plain C++17/OpenMP FP64
// restatement of the reference's self-consistency iteration with the
// reference's own loop structure: host-serial loop over Fermi angles, full
// coherence-function grids per angle, separate Green-function passes.  It is
// the checker for the HIP product and the CPU baseline of the benchmark.
//
// Citations are relative to the reference tree (superconga 1.1.0).
// Compile with -ffp-contract=off and without -march flags so that discrete
// geometric decisions are taken on plainly rounded FP64 arithmetic.

#include "conga_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

using cplx = std::complex<double>;
static thread_local std::string g_error;

// ---------------------------------------------------------------------------
// Constants -- include/conga/defines.h:25-46
// ---------------------------------------------------------------------------
constexpr int GRAIN_SIZE_MARGIN = 3;
constexpr double MAXIMUM_COORDINATE = 0.5;
constexpr double EULER_MASCHERONI = 0.57721566490153286060651209008240243;
static const double S_WAVE_BULK = 0.5 * std::exp(-EULER_MASCHERONI);
static const double D_WAVE_BULK =
    std::exp(-EULER_MASCHERONI - 0.5) / std::sqrt(2.0);

// Largest Ozaki pole for n+1 poles -- include/conga/Parameters.h:191-258.
// This table is data of the reference (a golden vector for the pole solver).
static const double kMaxOzakiPole[] = {
    3.464101615138,      13.043193723013,    27.182805923222,
    46.319508681820,     70.526709814647,    99.819209820145,
    134.201334576035,    173.674723843717,   218.240115243641,
    267.897882477874,    322.648232120917,   382.491286239281,
    447.427120870950,    517.455785454923,   592.577313298030,
    672.791727521291,    758.099044591524,   848.499276499387,
    943.992432147387,    1044.578518259219,  1150.257539991102,
    1261.029501351330,   1376.894405494319,  1497.852254930454,
    1623.903051677801,   1755.046797373544,  1891.283493357799,
    2032.613140735255,   2179.035740424027,  2330.551293191541,
    2487.159799683865,   2648.861260447426,  2815.655675948037,
    2987.543046583090,   3164.523372695167,  3346.596654578822,
    3533.762892490810,   3726.022086653727,  3923.374237262044,
    4125.819344486381,   4333.357408476837,  4545.988429367131,
    4763.712407272654,   4986.529342298375,  5214.439234539534,
    5447.442084074518,   5685.537890980710,  5928.726655323640,
    6177.008377162879,   6430.383056551208,  6688.850693540401,
    6952.411288172675,   7221.064840486898,  7494.811350521523,
    7773.650818306208,   8057.583243876886,  8346.608627256021,
    8640.726968470750,   8939.938267543248,  9244.242524494641,
    9553.639739344733,   9868.129912114642,  10187.713042813368,
    10512.389131458542,  10842.158178074462, 11177.020182659693,
    11516.975145232140,  11862.023065802561, 12212.163944387836,
    12567.397780986943,  12927.724575614844, 13293.144328276247,
    13663.657038985704,  14039.262707740532, 14419.961334558257,
    14805.752919440674,  15196.637462387956, 15592.614963412661,
    15993.685422515518,  16399.848839707749, 16811.105214993084,
    17227.454548365906,  17648.896839834808, 18075.432089412709,
    18507.060297090567,  18943.781462892035, 19385.595586789594,
    19832.502668811354,  20284.502708945201, 20741.595707203986,
    21203.781663584734,  21671.060578106339, 22143.432450728560,
    22620.897281488116,  23103.455070385800, 23591.105817420932,
    24083.849522596051,  24581.686185888688, 25084.615807335540,
    25592.638386922201,  26105.753924657907, 26623.962420530934,
    27147.263874548138,  27675.658286698381, 28209.145657019460,
    28747.725985486431,  29291.399272088380, 29840.165516850884,
    30394.024719756279,  30952.976880822123, 31517.022000048237,
    32086.160077425615,  32660.391112956720, 33239.715106652184,
    33824.132058499723,  34413.641968492826, 35008.244836653459,
    35607.940662971036,  36212.729447455276, 36822.611190103285,
    37437.585890915667,  38057.653549855168, 38682.814167005199,
    39313.067742276238,  39948.414275748975, 40588.853767362038,
    41234.386217142870,  41885.011625110244, 42540.729991207598,
    43201.541315517512,  43867.445597926955, 44538.442838554773,
    45214.533037321766,  45895.716194298584, 46581.992309429610,
    47273.361382657822,  47969.823414089085, 48671.378403742987,
    49378.026351488850,  50089.767257488224, 50806.601121605097,
    51528.527943908797,  52255.547724346325, 52987.660462994907,
    53724.866159820733,  54467.164814822114, 55214.556427974414,
    55967.040999261437,  56724.618528703817, 57487.289016365059,
    58255.052462227519,  59027.908866262900, 59805.858228478566,
    60588.900548765567,  61377.035827283733, 62170.264063961804,
    62968.585258870233,  63771.999411872166, 64580.506523058772,
    65394.106592470736,  66212.799620066056, 67036.585605816275,
    67865.464549688913,  68699.436451832487, 69538.501312039167,
    70382.659130497530,  71231.909907065638, 72086.253641840856,
    72945.690334811923,  73810.219985944655, 74679.842595227819,
    75554.558162766305,  76434.366688360693, 77319.268172226453,
    78209.262614181425,  79104.350014321433, 80004.530372739711,
    80909.803689260734,  81820.169963918539, 82735.629196799040,
    83656.181387807985,  84581.826537007699, 85512.564644378072,
    86448.395710174183,  87389.319733820099, 88335.336715730969,
    89286.446655858366,  90242.649554185831, 91203.945410558954,
    92170.334225223487,  93141.815998147023, 94118.390728990082,
    95100.058418232264,  96086.819065539064, 97078.672670988453,
    98075.619234544327,  99077.658756581528, 100084.791236587014,
    101097.016674872531, 102114.335071418158};
constexpr int kNumMaxOzakiPole = sizeof(kMaxOzakiPole) / sizeof(double);

// ---------------------------------------------------------------------------
// Small utilities -- include/conga/utils.h:34-183
// ---------------------------------------------------------------------------
static double wrapToRange(double v, double lo, double hi) {
  if (lo <= v && v < hi) return v;
  const double range = hi - lo;
  const double s = (v - lo) / range;
  return (s - std::floor(s)) * range + lo;
}
static int wrapToRangeInt(int v, int lo, int hi) {
  // utils.h:70-78 wraps through float arithmetic and rounds.
  const float value = static_cast<float>(v);
  const float flo = static_cast<float>(lo), fhi = static_cast<float>(hi);
  float r;
  if (flo <= value && value < fhi) {
    r = value;
  } else {
    const float range = fhi - flo;
    const float s = (value - flo) / range;
    r = (s - std::floor(s)) * range + flo;
  }
  return static_cast<int>(std::round(r));
}
static void rotate(double a, double x, double y, double &ox, double &oy) {
  const double c = std::cos(a), s = std::sin(a);
  ox = (x * c - y * s);
  oy = (x * s + y * c);
}
static int sgn(double v) { return int(0.0 < v) - int(v < 0.0); }

// Complex helpers written from the mathematics (no library complex division
// or square root is used for the arithmetic that matters).
// Principal square root, Kahan's formulation: with t = sqrt((|x|+|z|)/2),
// sqrt(z) = (t, y/2t) for x >= 0 and (|y|/2t, copysign(t, y)) otherwise.
static inline cplx csqrt(cplx z) {
  const double x = z.real(), y = z.imag();
  if (x == 0.0 && y == 0.0) return cplx(0.0, y);
  const double t = std::sqrt(0.5 * (std::fabs(x) + std::hypot(x, y)));
  if (x >= 0.0) return cplx(t, y / (2.0 * t));
  return cplx(std::fabs(y) / (2.0 * t), std::copysign(t, y));
}
// Smith's scaled complex division (Smith, CACM 1962).
static inline cplx cdiv(cplx a, cplx b) {
  const double c = b.real(), d = b.imag();
  if (std::fabs(c) >= std::fabs(d)) {
    const double r = d / c, den = c + d * r;
    return cplx((a.real() + a.imag() * r) / den,
                (a.imag() - a.real() * r) / den);
  }
  const double r = c / d, den = c * r + d;
  return cplx((a.real() * r + a.imag()) / den, (a.imag() * r - a.real()) / den);
}
static inline cplx cmul(cplx a, cplx b) {
  return cplx(a.real() * b.real() - a.imag() * b.imag(),
              a.real() * b.imag() + a.imag() * b.real());
}

// ---------------------------------------------------------------------------
// Riccati solutions -- include/conga/riccati/RiccatiSolutions.h:36-94
// ---------------------------------------------------------------------------
static void homogeneous(cplx delta, cplx energy, cplx &gamma, cplx &gtilde) {
  const double n = delta.real() * delta.real() + delta.imag() * delta.imag();
  const cplx omega = csqrt(cplx(n, 0.0) - cmul(energy, energy));
  const cplx iomega(-omega.imag(), omega.real());
  const cplx inv = cdiv(cplx(1.0, 0.0), energy + iomega);
  gamma = -cmul(delta, inv);
  gtilde = cmul(std::conj(delta), inv);
}
// Implicit midpoint step solved as a quadratic a y^2 + b y + c = 0.
static inline cplx riccatiStep(cplx delta, cplx energy, cplx y0, double h) {
  const cplx s(0.0, h);
  const cplx a = 0.25 * cmul(s, std::conj(delta));
  const cplx se = cmul(s, energy);
  const cplx b = 2.0 * cmul(a, y0) + se - 1.0;
  const cplx c = cmul(cmul(a, y0) + se + 1.0, y0) + cmul(delta, s);
  const cplx disc = csqrt(cmul(b, b) - 4.0 * cmul(a, c));
  return cdiv(2.0 * c, -b + disc);
}

// Green functions -- include/conga/green/LocalGreens.h:32-70
static inline cplx greensG(cplx g, cplx gt) {
  const cplx p = cmul(g, gt);
  const cplx tmp = -cdiv(1.0 - p, 1.0 + p);
  return cplx(-tmp.imag(), tmp.real());
}
static inline cplx greensDelta(cplx g, cplx gt) {
  const cplx tmp = cdiv(cplx(2.0, 0.0), 1.0 + cmul(g, gt));
  const cplx factor(-tmp.imag(), tmp.real());
  const cplx f = -cmul(factor, g);
  const cplx ftc = std::conj(cmul(factor, gt));
  return f + ftc;
}

// ---------------------------------------------------------------------------
// Parameters -- include/conga/Parameters.h:380-458
// ---------------------------------------------------------------------------
static int numOzakiEnergies(double T, double cutoff) {
  const double val = 2.0 * M_PI * (cutoff / T);
  const double *up =
      std::upper_bound(kMaxOzakiPole, kMaxOzakiPole + kNumMaxOzakiPole, val);
  if (up != kMaxOzakiPole + kNumMaxOzakiPole)
    return static_cast<int>(up - kMaxOzakiPole) + 1;
  const double c2 = 2.5465, c1 = 1.2744, c0 = 0.4651;
  int i = kNumMaxOzakiPole - 1;
  bool ok = false;
  while (!ok) {
    ++i;
    const double x = i;
    ok = x * (c2 * x + c1) + c0 >= val;
  }
  return i;
}
static void lattice(double width, double ppxi, int &n0, int &n1,
                    double &grainFraction) {
  const double frac = width * ppxi;
  const int total = static_cast<int>(std::ceil(frac)) + 2 * GRAIN_SIZE_MARGIN;
  grainFraction = frac / static_cast<double>(total);
  n0 = total;
  const double m = 0.5 * (std::sqrt(2.0) - 1.0) * static_cast<double>(total);
  n1 = total + 2 * static_cast<int>(std::ceil(m));
}

// ---------------------------------------------------------------------------
// Ozaki poles -- IntegrationIteratorOzaki.h:108-200.  The generalized problem
// A x = w B x (B = diag(2i+1), A tridiagonal with -1/2 off the diagonal) is
// reduced to C = B^-1/2 A B^-1/2 and diagonalized with the cyclic Jacobi
// method (Golub & Van Loan, Matrix Computations, sec. 8.5).  The B-normalized
// eigenvector x = B^-1/2 z has first component z_0.
// ---------------------------------------------------------------------------
static void ozakiPoles(int numPoles, std::vector<double> &poles,
                       std::vector<double> &residues) {
  const int N = 2 * numPoles;
  std::vector<double> C(N * N, 0.0), V(N * N, 0.0);
  for (int i = 0; i < N; ++i) V[i * N + i] = 1.0;
  for (int i = 0; i + 1 < N; ++i) {
    const double v =
        -0.5 / std::sqrt((2.0 * i + 1.0) * (2.0 * (i + 1) + 1.0));
    C[i * N + i + 1] = v;
    C[(i + 1) * N + i] = v;
  }
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < N; ++p)
      for (int q = p + 1; q < N; ++q) off += C[p * N + q] * C[p * N + q];
    if (std::sqrt(off) < 1e-18) break;
    for (int p = 0; p < N; ++p) {
      for (int q = p + 1; q < N; ++q) {
        const double apq = C[p * N + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double app = C[p * N + p], aqq = C[q * N + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) /
                         (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < N; ++k) {
          const double akp = C[k * N + p], akq = C[k * N + q];
          C[k * N + p] = c * akp - s * akq;
          C[k * N + q] = s * akp + c * akq;
        }
        for (int k = 0; k < N; ++k) {
          const double apk = C[p * N + k], aqk = C[q * N + k];
          C[p * N + k] = c * apk - s * aqk;
          C[q * N + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < N; ++k) {
          const double vkp = V[k * N + p], vkq = V[k * N + q];
          V[k * N + p] = c * vkp - s * vkq;
          V[k * N + q] = s * vkp + c * vkq;
        }
      }
    }
  }
  std::vector<int> order(N);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(),
            [&](int a, int b) { return C[a * N + a] < C[b * N + b]; });
  poles.resize(numPoles);
  residues.resize(numPoles);
  for (int i = 0; i < numPoles; ++i) {
    const int j = order[i];
    const double w = C[j * N + j];
    const double pole = -1.0 / w;
    const double x0 = V[0 * N + j]; // B_00 = 1
    const double rs = 0.5 * x0 * pole;
    poles[i] = pole;
    residues[i] = rs * rs;
  }
}

// ---------------------------------------------------------------------------
// Bulk order parameter -- include/conga/order_parameter/OrderParameterBulk.h
// ---------------------------------------------------------------------------
static double bulkResidual(double d, double T, int n, double g) {
  double s = 0.0;
  for (int k = n - 1; k >= 0; --k) {
    const double e = T * (static_cast<double>(k) + 0.5);
    s += 1.0 / std::hypot(d, e);
  }
  return d * (s * g - 1.0);
}
static double bulkGap(double T) {
  const double crit = 10.0 * std::numeric_limits<double>::epsilon();
  const int maxIter = 10;
  const double cutoff = 100.0;
  if (T >= 1.0) return 0.0;
  if (T < std::numeric_limits<double>::epsilon()) return S_WAVE_BULK;
  const int n = static_cast<int>(cutoff / T - 0.5) + 1;
  double es = 0.0;
  for (int k = n - 1; k >= 0; --k) es += 1.0 / (T * (static_cast<double>(k) + 0.5));
  const double g = 1.0 / (std::log(T) / T + es);
  double op = S_WAVE_BULK * std::sqrt(1.0 - T * T * T);
  double x[2] = {op, std::sqrt(op)};
  double r[2];
  for (int i = 0; i < 2; ++i) r[i] = bulkResidual(x[i], T, n, g);
  for (int it = 0; it < maxIter; ++it) {
    const double opRatio = x[0] / x[1];
    const double resRatio = r[0] / r[1];
    const double w = 1.0 / (1.0 - opRatio / resRatio);
    op = (1.0 - w) * x[0] + w * x[1];
    const double res = bulkResidual(op, T, n, g);
    const bool small = op < 10.0 * std::numeric_limits<double>::epsilon();
    const double rel = small ? res : std::fabs(res / op);
    if (rel < crit) break;
    const int idx = std::fabs(r[0]) > std::fabs(r[1]) ? 0 : 1;
    x[idx] = op;
    r[idx] = res;
  }
  return op;
}
static double bulkFreeEnergy(double T, double d) {
  const double cutoff = 100.0;
  const int n = static_cast<int>(cutoff / T - 0.5) + 1;
  double s = 0.0;
  for (int k = n - 1; k >= 0; --k) {
    const double e = T * (static_cast<double>(k) + 0.5);
    const double om = std::hypot(d, e);
    s += (1.0 / e - 2.0 / (e + om));
  }
  return d * d * (std::log(T) + T * s);
}
static double yoshida(double T, double d, double cutoff) {
  if (T < std::numeric_limits<double>::epsilon()) return 1.0;
  const int n = static_cast<int>(cutoff / T - 0.5) + 1;
  double s = 0.0;
  for (int k = n - 1; k >= 0; --k) {
    const double e = T * (static_cast<double>(k) + 0.5);
    const double om = std::hypot(d, e);
    s += 1.0 / std::pow(om, 3);
  }
  return T * d * d * s;
}

// ---------------------------------------------------------------------------
// Geometry -- include/conga/geometry/{Disc,Polygon,GeometryComponent,
// GeometryGroup,geometry_utils}.h
// ---------------------------------------------------------------------------
struct Component {
  int type = OC_DISC; // OC_DISC or polygon (regular polygons become vertices)
  bool add = true;
  double radius = 0.5, cx = 0.0, cy = 0.0; // disc, internal units
  std::vector<double> vx, vy;              // polygon, internal units
};

struct FrameGeometry {
  std::vector<signed char> label;
  std::vector<double> nx, ny, dist;
};

// DiscGeometry.h:82-173 -- 3x3 row-majority label.
static signed char discLabel(int X, int Y, double R, double ox, double oy) {
  const double r2 = R * R;
  bool nb[3];
  for (int j = 0; j < 3; ++j) {
    const double ny = static_cast<double>(Y - 1 + j) - oy;
    int rowSum = 0;
    for (int i = 0; i < 3; ++i) {
      const double nx = static_cast<double>(X - 1 + i) - ox;
      rowSum += (nx * nx + ny * ny <= r2) ? 1 : 0;
    }
    nb[j] = rowSum > 1;
  }
  if (nb[0] && nb[1] && nb[2]) return 1;
  if (!nb[0] && nb[1] && nb[2]) return 2;
  if (nb[0] && nb[1] && !nb[2]) return 3;
  return 0;
}
// DiscGeometry.h:66-75
static double convertCoordinate(double f, int nSim, int nGrid) {
  const int diff = nGrid - nSim;
  const double half = 0.5 * static_cast<double>(diff);
  return static_cast<double>(nSim - 1) * (f + MAXIMUM_COORDINATE) + half;
}

// DiscGeometry.h:175-221,269-294
static void createDisc(const Component &c, double angle, int N0, int N1,
                       FrameGeometry &g) {
  double rcx, rcy;
  rotate(angle, c.cx, c.cy, rcx, rcy);
  const double ox = convertCoordinate(rcx, N0, N1);
  const double oy = convertCoordinate(rcy, N0, N1);
  const double R = c.radius * static_cast<double>(N0 - 1);
  g.label.assign(size_t(N1) * N1, 0);
  g.nx.assign(size_t(N1) * N1, 0.0);
  g.ny.assign(size_t(N1) * N1, 0.0);
  g.dist.assign(size_t(N1) * N1, 0.0);
  for (int ty = 0; ty < N1; ++ty)
    for (int tx = 0; tx < N1; ++tx) {
      const signed char lab = discLabel(tx, ty, R, ox, oy);
      const double x = static_cast<double>(tx) - ox;
      const double y = static_cast<double>(ty) - oy;
      double d = 0.0, nxv = 0.0, nyv = 0.0;
      if (lab > 1) {
        d = std::sqrt(R * R - x * x) - std::fabs(y);
        nxv = -x / R;
        nyv = (y <= 0.0) ? -(y - d) / R : -(y + d) / R;
      }
      const size_t idx = size_t(ty) * N1 + tx;
      g.label[idx] = lab;
      g.dist[idx] = d;
      g.nx[idx] = nxv;
      g.ny[idx] = nyv;
    }
}

// PolygonGeometry.h:101-190 + GeometryComponent.h:112-147 (entry/exit
// labelling) + PolygonGeometry.h:233-251 (masking).
static void createPolygon(const Component &c, double angle, int N0, int N1,
                          FrameGeometry &g) {
  const int nv = static_cast<int>(c.vx.size());
  const double scale = static_cast<double>(N0 - 1);
  std::vector<signed char> inside(size_t(N1) * N1, 0);
  g.nx.assign(size_t(N1) * N1, 0.0);
  g.ny.assign(size_t(N1) * N1, 0.0);
  g.dist.assign(size_t(N1) * N1, 0.0);
  const double ca = std::cos(angle), sa = std::sin(angle);
  std::vector<double> px(nv), py(nv);
  for (int i = 0; i < nv; ++i) {
    px[i] = (ca * c.vx[i] - sa * c.vy[i]) * scale;
    py[i] = (sa * c.vx[i] + ca * c.vy[i]) * scale;
  }
  for (int ty = 0; ty < N1; ++ty)
    for (int tx = 0; tx < N1; ++tx) {
      const double X = (double)tx - (double)(N1 - 1) * 0.5;
      const double Y = (double)ty - (double)(N1 - 1) * 0.5;
      int sumCross = 0;
      double bd = 100000.0, nxv = 0.0, nyv = 0.0;
      for (int i = 0; i < nv; ++i) {
        const int ip = (i + 1 == nv) ? 0 : i + 1;
        const double lx = px[ip] - px[i], ly = py[ip] - py[i];
        const double cz = lx * (Y - py[i]) - ly * (X - px[i]);
        if (cz >= 0.0) sumCross++;
        double tnx = -ly, tny = lx;
        const double nrm = 1.0 / std::sqrt(tnx * tnx + tny * tny);
        tnx *= nrm;
        tny *= nrm;
        double tb;
        if (std::fabs(lx) > 0.00001) {
          const double k = ly / lx;
          const double m = py[i] - k * px[i];
          tb = std::fabs(k * X + m - Y);
        } else {
          tb = 10000.0;
        }
        if (tb < bd) {
          bd = tb;
          nxv = tnx;
          nyv = tny;
        }
      }
      const size_t idx = size_t(ty) * N1 + tx;
      inside[idx] = (sumCross == nv) ? 1 : 0;
      g.dist[idx] = bd;
      g.nx[idx] = nxv;
      g.ny[idx] = nyv;
    }
  // labelEntryExit_Kernel
  g.label.assign(size_t(N1) * N1, 0);
  for (int ty = 0; ty < N1; ++ty)
    for (int tx = 0; tx < N1; ++tx) {
      signed char nl = 0;
      if (ty > 0 && ty < N1 - 1) {
        const signed char o0 = inside[size_t(ty - 1) * N1 + tx];
        const signed char o1 = inside[size_t(ty) * N1 + tx];
        const signed char o2 = inside[size_t(ty + 1) * N1 + tx];
        if (o1 == 1) {
          nl = 1;
          if (o0 == 0)
            nl = 2;
          else if (o2 == 0)
            nl = 3;
          if (o0 == 0 && o2 == 0) nl = 0;
        }
      }
      g.label[size_t(ty) * N1 + tx] = nl;
    }
  // maskNormals_Kernel
  for (size_t i = 0; i < g.label.size(); ++i) {
    const double mult = (int(g.label[i]) > 1) ? 1.0 : 0.0;
    g.dist[i] *= mult;
    g.nx[i] *= mult;
    g.ny[i] *= mult;
  }
}

// GeometryGroup.h:487-497 operation tables.
static const signed char kOpAdd[16] = {0, 1, 2, 3, 1, 1, 1, 1,
                                       2, 1, 2, 1, 3, 1, 1, 3};

// GeometryGroup.h:499-522 + geomOp_kernel (216-254) + geomOpRemove_kernel
// (272-319).
static void createGroup(const std::vector<Component> &comps, double angle,
                        int N0, int N1, FrameGeometry &out) {
  const size_t n = size_t(N1) * N1;
  out.label.assign(n, 0);
  out.nx.assign(n, 0.0);
  out.ny.assign(n, 0.0);
  out.dist.assign(n, 0.0);
  FrameGeometry c;
  for (const Component &comp : comps) {
    if (comp.type == OC_DISC)
      createDisc(comp, angle, N0, N1, c);
    else
      createPolygon(comp, angle, N0, N1, c);
    if (comp.add) {
      for (size_t i = 0; i < n; ++i) {
        signed char d0 = out.label[i];
        const signed char d1 = c.label[i];
        const double dist0 = out.dist[i], dist1 = c.dist[i];
        d0 = kOpAdd[d0 * 4 + d1];
        double nxv, nyv;
        if (d0 == 2 || d0 == 3) {
          nxv = out.nx[i] + c.nx[i];
          nyv = out.ny[i] + c.ny[i];
          const double nrm = 1.0 / std::sqrt(nxv * nxv + nyv * nyv);
          nxv *= nrm;
          nyv *= nrm;
        } else {
          nxv = 0.0;
          nyv = 0.0;
        }
        out.nx[i] = nxv;
        out.ny[i] = nyv;
        out.label[i] = d0;
        out.dist[i] = std::fmax(dist0, dist1);
      }
    } else {
      std::vector<signed char> tmp(n, 0);
      for (int ty = 1; ty < N1 - 1; ++ty)
        for (int tx = 0; tx < N1; ++tx) {
          const size_t im = size_t(ty - 1) * N1 + tx, i = size_t(ty) * N1 + tx,
                       ip = size_t(ty + 1) * N1 + tx;
          signed char d0 = out.label[i];
          if (c.label[i] > 0) {
            d0 = 0;
          } else if (c.label[ip] == 2) {
            if (d0 == 1) {
              d0 = 3;
              double nxv = out.nx[i] - c.nx[ip], nyv = out.ny[i] - c.ny[ip];
              const double nrm = 1.0 / std::sqrt(nxv * nxv + nyv * nyv);
              out.nx[i] = nxv * nrm;
              out.ny[i] = nyv * nrm;
              out.dist[i] = 1.0 - c.dist[ip];
            }
          } else if (c.label[im] == 3) {
            if (d0 == 1) {
              d0 = 2;
              double nxv = out.nx[i] - c.nx[im], nyv = out.ny[i] - c.ny[im];
              const double nrm = 1.0 / std::sqrt(nxv * nxv + nyv * nyv);
              out.nx[i] = nxv * nrm;
              out.ny[i] = nyv * nrm;
              out.dist[i] = 1.0 - c.dist[im];
            }
          }
          tmp[i] = d0;
        }
      out.label.swap(tmp);
    }
  }
}

// GeometryGroup.h:524-546, DiscGeometry.h:223-242,296-323,
// PolygonGeometry.h:192-231,383-402 -- reference-frame domain on N0.
static std::vector<signed char> createDomain(const std::vector<Component> &comps,
                                             int N0) {
  const size_t n = size_t(N0) * N0;
  std::vector<signed char> dom(n, 0), loc(n, 0);
  for (const Component &comp : comps) {
    if (comp.type == OC_DISC) {
      const double ox = convertCoordinate(comp.cx, N0, N0);
      const double oy = convertCoordinate(comp.cy, N0, N0);
      const double R = comp.radius * static_cast<double>(N0 - 1);
      for (int ty = 0; ty < N0; ++ty)
        for (int tx = 0; tx < N0; ++tx)
          loc[size_t(ty) * N0 + tx] = discLabel(tx, ty, R, ox, oy) > 0 ? 1 : 0;
    } else {
      const int nv = static_cast<int>(comp.vx.size());
      for (int ty = 0; ty < N0; ++ty)
        for (int tx = 0; tx < N0; ++tx) {
          const double X =
              static_cast<double>(tx) / static_cast<double>(N0 - 1) - 0.5;
          const double Y =
              static_cast<double>(ty) / static_cast<double>(N0 - 1) - 0.5;
          int sc = 0;
          for (int i = 0; i < nv; ++i) {
            const int ip = (i == nv - 1) ? 0 : i + 1;
            const double lx = comp.vx[ip] - comp.vx[i];
            const double ly = comp.vy[ip] - comp.vy[i];
            const double cz = lx * (Y - comp.vy[i]) - ly * (X - comp.vx[i]);
            if (cz >= 0.0) sc++;
          }
          loc[size_t(ty) * N0 + tx] = (sc == nv) ? 1 : 0;
        }
    }
    for (size_t i = 0; i < n; ++i) {
      if (comp.add) {
        if (loc[i]) dom[i] = 1;
      } else {
        const bool v0 = dom[i] > 0, v1 = loc[i] > 0;
        dom[i] = (!v0 || v1) ? 0 : 1;
      }
    }
  }
  return dom;
}

// Polygon.h:536-561 -- vertices of a regular polygon.
static void regularPolygonVertices(int n, double radius, double angle, double cx,
                                   double cy, std::vector<double> &vx,
                                   std::vector<double> &vy) {
  vx.resize(n);
  vy.resize(n);
  const double step = 2.0 * M_PI / static_cast<double>(n);
  const double off = step * 0.5 + angle;
  for (int i = 0; i < n; ++i) {
    const double a = step * static_cast<double>(i) + off;
    vx[i] = radius * std::cos(a) + cx;
    vy[i] = radius * std::sin(a) + cy;
  }
}

// ---------------------------------------------------------------------------
// Grid helpers (planar complex: [re N^2][im N^2])
// ---------------------------------------------------------------------------
// GridFunctions_GPU.h:436-544.  The rotation angle is a float and its
// cosine/sine are evaluated in single precision; the bilinear interpolation is
// double.  'additive' selects insertRotatedGridAdditive.
static void insertRotated(double *dst, int Nx, int Ny, const double *src,
                          int Nxo, int Nyo, float phi, bool additive) {
  const double xt = (Nxo - Nx) / 2.0, yt = (Nyo - Ny) / 2.0;
  const double sx = (Nxo - 1) / 2.0, sy = (Nyo - 1) / 2.0;
  const double c = static_cast<double>(std::cos(-phi));
  const double s = static_cast<double>(std::sin(-phi));
  for (int y = 0; y < Ny; ++y)
    for (int x = 0; x < Nx; ++x) {
      const double X = x + xt - sx;
      const double Y = y + yt - sy;
      const double Xr = X * c - Y * s + sx;
      const double Yr = X * s + Y * c + sy;
      const int Xi = static_cast<int>(Xr), Yi = static_cast<int>(Yr);
      if (Xi >= 0 && Xi + 1 < Nxo && Yi >= 0 && Yi + 1 < Nyo) {
        const double fx = Xr - (double)Xi, fy = Yr - (double)Yi;
        const double a = (1.0 - fx) * src[size_t(Yi) * Nxo + Xi] +
                         fx * src[size_t(Yi) * Nxo + Xi + 1];
        const double b = (1.0 - fx) * src[size_t(Yi + 1) * Nxo + Xi] +
                         fx * src[size_t(Yi + 1) * Nxo + Xi + 1];
        const double v = (1.0 - fy) * a + fy * b;
        if (additive)
          dst[size_t(y) * Nx + x] += v;
        else
          dst[size_t(y) * Nx + x] = v;
      }
    }
}
// domain_ops.h:11-57 -- extend values outwards from the boundary (in place,
// per plane).  Only label-0 sites are written and only label>0 sites are
// read, so the in-place update is order independent.
static void extendValues(const signed char *labels, int N, double *plane) {
  const int ext = 2;
  for (int y = 0; y < N; ++y)
    for (int x = 0; x < N; ++x) {
      if (labels[size_t(y) * N + x] != 0) continue;
      const int x0 = std::max(x - ext, 0), x1 = std::min(x + ext, N - 1);
      const int y0 = std::max(y - ext, 0), y1 = std::min(y + ext, N - 1);
      double val = 0.0;
      int np = 0, best = 1000;
      for (int ky = y0; ky <= y1; ++ky)
        for (int kx = x0; kx <= x1; ++kx) {
          if (labels[size_t(ky) * N + kx] > 0) {
            const int dx = x - kx, dy = y - ky, d2 = dx * dx + dy * dy;
            const double v = plane[size_t(ky) * N + kx];
            if (d2 < best) {
              val = v;
              np = 1;
              best = d2;
            } else if (d2 == best) {
              val += v;
              np++;
            }
          }
        }
      if (np > 0) plane[size_t(y) * N + x] = val / static_cast<double>(np);
    }
}

// ComputeResidual.h:54-106 -- residual of old/new arrays.
static double residualArrays(const double *oldv, const double *newv, size_t n,
                             int norm) {
  double numer = 0.0, denom = 0.0, eps = 1e-4;
  if (norm == OC_L1) {
    for (size_t i = 0; i < n; ++i) {
      numer += std::fabs(oldv[i] - newv[i]);
      denom += std::fabs(oldv[i]);
    }
    eps *= static_cast<double>(n);
  } else if (norm == OC_L2) {
    for (size_t i = 0; i < n; ++i) {
      const double d = oldv[i] - newv[i];
      numer += d * d;
      denom += oldv[i] * oldv[i];
    }
    numer = std::sqrt(numer);
    denom = std::sqrt(denom);
    eps *= static_cast<double>(n);
  } else {
    // LInfNorm (GridData.h): max(|min(v, DBL_MAX)|, |max(v, DBL_MIN)|).
    auto linf = [&](auto get) {
      double mn = std::numeric_limits<double>::max();
      double mx = std::numeric_limits<double>::min();
      for (size_t i = 0; i < n; ++i) {
        const double v = get(i);
        mn = std::min(mn, v);
        mx = std::max(mx, v);
      }
      return std::max(std::fabs(mn), std::fabs(mx));
    };
    numer = linf([&](size_t i) { return oldv[i] - newv[i]; });
    denom = linf([&](size_t i) { return oldv[i]; });
  }
  return numer / std::max(denom, eps);
}
static double residualScalar(double oldv, double newv) {
  return std::fabs(oldv - newv) / std::max(std::fabs(oldv), 1e-4);
}

// ---------------------------------------------------------------------------
// The simulation state.
// ---------------------------------------------------------------------------
struct Oracle {
  oc_config cfg{};
  int N0 = 0, N1 = 0, Ne = 0, Na = 0, Nc = 0;
  double h = 0, grainFraction = 0, T = 0, lambda = -1;
  int numGrain = 0;
  std::vector<Component> comps;
  std::vector<signed char> domainRef;
  std::vector<FrameGeometry> frames; // per angle (precomputed once; the
                                     // reference recreates the identical
                                     // frame on every use)
  // Fermi surface
  std::vector<double> systemAngle, polar;
  // energies: eps_e = energyRe[e] + i energyIm[e] (energyRe is zero for the
  // Ozaki poles of the self-consistency)
  std::vector<double> energyRe, energyIm, residue;
  // Keldysh spectroscopy (IntegrationIteratorKeldysh.h, ComputeLDOS.h)
  struct Spectroscopy {
    bool on = false;
    double eMin = 0, eMax = 0, broadening = 0;
    int numEnergies = 0, perBlock = 0;
    int blockStart = 0, lastBlock = -1;
    std::vector<double> all;      // numEnergies x N0^2
    std::vector<double> interval; // Ne x N0^2
    double residual = 0;
    bool partial = false, converged = false;
  } kel;
  // order parameter
  std::vector<int> symmetry;
  std::vector<double> tc, basisNorm, coupling;
  std::vector<std::vector<double>> delta;    // context x (2*N0^2 each)
  std::vector<std::vector<double>> deltaNew; // compute object result
  std::vector<double> A, Anew;               // 2*N0^2 (scaled by lambda^2)
  // boundary storage
  std::vector<int> idxInG, idxOutG, idxInGt, idxOutGt, numIsG, numIsGt;
  std::vector<int> coordG, coordGt;
  std::vector<int> xmin, xmax, ymin, ymax;
  long long nInG = 0, nOutG = 0, nInGt = 0, nOutGt = 0;
  std::vector<double> bInG, bOutG, bInGt, bOutGt; // planar 2*count
  bool boundaryInit = false;
  // coherence functions (planar per energy: e*2*N1^2 + y*N1 + x)
  std::vector<double> gamma, gammat;
  // compute-object state
  std::vector<double> current, currentAcc; // current 2*N0^2 real; acc complex
  std::vector<double> feGrid;              // complex N0^2
  double feTotal = 0, fePerArea = 0, feMag = 0, feMagPerArea = 0;
  bool firstCurrent = true, firstFE = true, firstCompute = true;
  double resBoundary = std::numeric_limits<double>::max();
  double resOP = 0, resJ = 0, resFE = 0, resA = 0;
  bool convOP = false, convJ = false, convFE = false;
  int numIterations = 0;
  // vector potential
  std::vector<double> filter2D;
  // accelerator state
  int accIter = 0;
  std::vector<std::vector<double>> accX, accD;
  std::vector<double> accSq;
  std::vector<size_t> accIdx;
  std::vector<double> accMean, accStep, accInertia, accPrev, accVel;
  double bbStep = 0;
  bool accInit = false;
  // Broyden: pairs (c_i, d_i) of H = -alpha I + sum c_i d_i^T, previous x, F
  std::vector<std::vector<double>> brC, brD;
  std::vector<double> brLastX, brLastF;

  int numEnergies() const { return Ne; }
  size_t n0sq() const { return size_t(N0) * N0; }
  size_t n1sq() const { return size_t(N1) * N1; }
};

// FermiSurface.h:119-174, FermiSurfaceCircle.h:83-163
static double polarAngle(int k, int N) {
  const double u = static_cast<double>(M_PI) * static_cast<double>(2 * k) /
                       static_cast<double>(N) +
                   static_cast<double>(0.5 * M_PI);
  return wrapToRange(u, 0.0, 2.0 * M_PI);
}
static double angleWeight(int k, int N) {
  const int w = wrapToRangeInt(k, 0, N);
  if (N % 4 == 0) {
    if (w % 4 == 0) return 28.0 / 45.0;
    if (w % 2 == 0) return 24.0 / 45.0;
    return 64.0 / 45.0;
  } else if (N % 3 == 0) {
    return (w % 3 == 0) ? 6.0 / 8.0 : 9.0 / 8.0;
  } else if (N % 2 == 0) {
    return (w % 2 == 0) ? 2.0 / 3.0 : 4.0 / 3.0;
  }
  return 1.0;
}
static double fractionalAngleIndex(double px, double py, int N) {
  const double kx = wrapToRange(px, 0.0 - static_cast<double>(M_PI),
                                0.0 + static_cast<double>(M_PI));
  const double ky = wrapToRange(py, 0.0 - static_cast<double>(M_PI),
                                0.0 + static_cast<double>(M_PI));
  const double u = std::atan2(ky - 0.0, kx - 0.0) - static_cast<double>(0.5 * M_PI);
  const double phi = wrapToRange(u, 0.0, 2.0 * M_PI);
  return static_cast<double>(N) * phi / static_cast<double>(2.0 * M_PI);
}

// OrderParameterBasis.h:144-221 (circular Fermi surface)
static double basisRaw(int sym, double px, double py) {
  switch (sym) {
  case OC_SWAVE: return 1.0;
  case OC_DX2Y2: return px * px - py * py;
  case OC_DXY: return px * py;
  case OC_GWAVE: return px * py * (px * px - py * py);
  }
  return 0.0;
}
static double basisEval(const Oracle &o, int c, double px, double py) {
  return basisRaw(o.symmetry[c], px, py) * o.basisNorm[c];
}
// OrderParameterComponent.h:197-208: evaluated at the momentum rotated by the
// crystal-axes rotation.
static double basisMomentum(const Oracle &o, int c, int k) {
  const double phi = polarAngle(k, o.Na);
  double rx, ry;
  rotate(o.cfg.crystal_axes_rotation, std::cos(phi), std::sin(phi), rx, ry);
  return basisEval(o, c, rx, ry);
}

// BoundaryStorage.h:298-527 + GeometryGroup.h:750-894
static void initBoundaryStorage(Oracle &o) {
  const int N1 = o.N1, Na = o.Na, Ne = o.Ne;
  o.idxInG.assign(size_t(Na) * N1, 0);
  o.idxOutG = o.idxInG;
  o.idxInGt = o.idxInG;
  o.idxOutGt = o.idxInG;
  o.numIsG = o.idxInG;
  o.numIsGt = o.idxInG;
  long long sInG = 0, sOutG = 0, sInGt = 0, sOutGt = 0;
  for (int k = 0; k < Na; ++k) {
    const signed char *dom = o.frames[k].label.data();
    // gamma: along +y
    for (int i = 0; i < N1; ++i) {
      int nin = 0, nout = 0;
      for (int j = 0; j < N1; ++j) {
        const signed char l = dom[size_t(j) * N1 + i];
        if (l == 2) nin++;
        if (l == 3) nout++;
      }
      o.numIsG[size_t(k) * N1 + i] = nout;
      o.idxInG[size_t(k) * N1 + i] = nin == 0 ? -1 : int(sInG);
      sInG += nin;
      o.idxOutG[size_t(k) * N1 + i] = nout == 0 ? -1 : int(sOutG);
      sOutG += nout;
    }
    // gamma tilde: along -y (entry 3, exit 2)
    for (int i = 0; i < N1; ++i) {
      int nin = 0, nout = 0;
      for (int j = N1 - 1; j >= 0; --j) {
        const signed char l = dom[size_t(j) * N1 + i];
        if (l == 3) nin++;
        if (l == 2) nout++;
      }
      o.numIsGt[size_t(k) * N1 + i] = nout;
      o.idxInGt[size_t(k) * N1 + i] = nin == 0 ? -1 : int(sInGt);
      sInGt += nin;
      o.idxOutGt[size_t(k) * N1 + i] = nout == 0 ? -1 : int(sOutGt);
      sOutGt += nout;
    }
  }
  o.nInG = sInG;
  o.nOutG = sOutG;
  o.nInGt = sInGt;
  o.nOutGt = sOutGt;
  o.coordG.clear();
  o.coordGt.clear();
  for (int k = 0; k < Na; ++k) {
    const signed char *dom = o.frames[k].label.data();
    for (int i = 0; i < N1; ++i)
      for (int j = 0; j < N1; ++j)
        if (dom[size_t(j) * N1 + i] == 3) o.coordG.push_back(j);
    for (int i = 0; i < N1; ++i)
      for (int j = N1 - 1; j >= 0; --j)
        if (dom[size_t(j) * N1 + i] == 2) o.coordGt.push_back(j);
  }
  o.bInG.assign(size_t(2 * sInG * Ne), 0.0);
  o.bOutG.assign(size_t(2 * sOutG * Ne), 0.0);
  o.bInGt.assign(size_t(2 * sInGt * Ne), 0.0);
  o.bOutGt.assign(size_t(2 * sOutGt * Ne), 0.0);
  // computeIntersectingTrajectories
  o.xmin.assign(Na, 0);
  o.xmax.assign(Na, 0);
  o.ymin.assign(Na, 0);
  o.ymax.assign(Na, 0);
  for (int k = 0; k < Na; ++k) {
    const int *sp = &o.idxInG[size_t(k) * N1];
    o.xmin[k] = 0;
    for (int i = 0; i < N1; ++i)
      if (sp[i] > -1) {
        o.xmin[k] = i;
        break;
      }
    o.xmax[k] = o.xmin[k];
    for (int i = N1 - 1; i >= o.xmin[k]; --i)
      if (sp[i] > -1) {
        o.xmax[k] = i;
        break;
      }
    const signed char *dom = o.frames[k].label.data();
    o.ymin[k] = N1 - 1;
    o.ymax[k] = 0;
    for (int i = o.xmin[k]; i <= o.xmax[k]; ++i) {
      for (int j = 0; j < N1; ++j)
        if (dom[size_t(j) * N1 + i] > 0) {
          if (j < o.ymin[k]) {
            o.ymin[k] = j;
            break;
          }
        }
      for (int j = N1 - 1; j >= 0; --j)
        if (dom[size_t(j) * N1 + i] > 0) {
          if (j > o.ymax[k]) {
            o.ymax[k] = j;
            break;
          }
        }
    }
  }
}

// Gauges.h:25-62
static void gaugeA(int gauge, double B, double area, double x, double y,
                   double &ax, double &ay) {
  if (gauge == OC_LANDAU) {
    ax = 0.0;
    ay = x * B;
  } else if (gauge == OC_SYMMETRIC) {
    const double f = 0.5 * B;
    ax = -y * f;
    ay = x * f;
  } else {
    const double r2 = std::max(x * x + y * y, 1e-7);
    const double f = B * area / (static_cast<double>(2.0 * M_PI) * r2);
    ax = -y * f;
    ay = x * f;
  }
}

// RiccatiSolver.h:127-176 + VectorPotential_Kernel.h:21-99
static void energyShift(const Oracle &o, int k, std::vector<double> &es) {
  const int N1 = o.N1;
  es.assign(o.n1sq(), 0.0);
  const double area = o.h * o.h * static_cast<double>(o.numGrain);
  const double B = static_cast<double>(M_PI) * o.cfg.external_flux_quanta / area;
  const double sa = o.systemAngle[k];
  const double phi = polarAngle(k, o.Na);
  const double vx = std::cos(phi), vy = std::sin(phi);
  const int q = sgn(o.cfg.charge_sign);
  std::vector<double> Ar;
  const bool useInduced = o.lambda > 0.0;
  if (!useInduced && !(std::fabs(B) > 0.0)) return;
  if (useInduced) {
    Ar.assign(2 * o.n1sq(), 0.0);
    for (int f = 0; f < 2; ++f)
      insertRotated(Ar.data() + f * o.n1sq(), N1, N1, o.A.data() + f * o.n0sq(),
                    o.N0, o.N0, static_cast<float>(-sa), false);
    const double sc = 1.0 / (o.lambda * o.lambda);
    for (double &v : Ar) v *= sc;
  }
  const double off = -0.5 * static_cast<double>(N1 - 1);
  for (int yi = 0; yi < N1; ++yi)
    for (int xi = 0; xi < N1; ++xi) {
      double x = (static_cast<double>(xi) + off) * o.h;
      double y = (static_cast<double>(yi) + off) * o.h;
      double rx, ry;
      rotate(sa, x, y, rx, ry);
      double ax, ay;
      gaugeA(o.cfg.gauge, B, area, rx, ry, ax, ay);
      const size_t idx = size_t(yi) * N1 + xi;
      if (useInduced) {
        ax = Ar[idx] + ax;
        ay = Ar[o.n1sq() + idx] + ay;
      }
      es[idx] = static_cast<double>(q) * (vx * ax + vy * ay);
    }
}

// OrderParameter.h:232-269 -- rotated total order parameter for angle k.
static void rotatedOrderParameter(const Oracle &o, int k,
                                  std::vector<double> &out) {
  const size_t n0 = o.n0sq();
  std::vector<double> sum(2 * n0, 0.0);
  // momentum of index k (the reference calls getRotatedGrid(-k) which then
  // evaluates the momentum of -(-k)).
  for (int c = 0; c < o.Nc; ++c) {
    const double b = basisMomentum(o, c, k);
    const double *d = o.delta[c].data();
    for (size_t i = 0; i < n0; ++i) {
      // complex * (b, 0): re*b - im*0, re*0 + im*b
      sum[i] += d[i] * b - d[n0 + i] * 0.0;
      sum[n0 + i] += d[i] * 0.0 + d[n0 + i] * b;
    }
  }
  extendValues(o.domainRef.data(), o.N0, sum.data());
  extendValues(o.domainRef.data(), o.N0, sum.data() + n0);
  const int kk = wrapToRangeInt(-k, 0, o.Na);
  const float phi = static_cast<float>(o.systemAngle[kk]);
  out.assign(2 * o.n1sq(), 0.0);
  insertRotated(out.data(), o.N1, o.N1, sum.data(), o.N0, o.N0, phi, true);
  insertRotated(out.data() + o.n1sq(), o.N1, o.N1, sum.data() + n0, o.N0, o.N0,
                phi, true);
}

static inline size_t sparseIdx(const int *base, int N1, int k, int x, int cnt,
                               int Ne, int e) {
  return size_t((base[size_t(k) * N1 + x] + cnt) * Ne + e);
}

// RiccatiSolverConfined.h:129-225 -- one site of the trajectory.
template <bool GAMMA>
static inline void siteStep(const Oracle &o, double step, int k, int e, int x,
                            int y, cplx energy, cplx energyPrev, cplx delta,
                            cplx deltaPrev, signed char label, const double *dist,
                            const int *idxIn, const int *idxOut,
                            const double *bin, long long nin, double *bout,
                            long long nout, int &cntIn, int &cntOut, cplx &coh) {
  const int ENTRY = GAMMA ? 2 : 3, EXIT = GAMMA ? 3 : 2;
  const int N1 = o.N1, Ne = o.Ne;
  const size_t sidx = size_t(y) * N1 + x;
  if (label == 1) {
    coh = riccatiStep((deltaPrev + delta) * 0.5, (energyPrev + energy) * 0.5,
                      coh, step);
  } else if (label == ENTRY) {
    const size_t sp = sparseIdx(idxIn, N1, k, x, cntIn, Ne, e);
    const cplx bc(bin[sp], bin[size_t(nin * Ne) + sp]);
    const double d = std::fabs(dist[sidx]);
    coh = riccatiStep(delta, energy, bc, d * step);
    cntIn++;
  } else if (label == EXIT) {
    coh = riccatiStep((deltaPrev + delta) * 0.5, (energyPrev + energy) * 0.5,
                      coh, step);
    const double d = std::fabs(dist[sidx]);
    const cplx bv = riccatiStep(delta, energy, coh, d * step);
    const size_t sp = sparseIdx(idxOut, N1, k, x, cntOut, Ne, e);
    bout[sp] = bv.real();
    bout[size_t(nout * Ne) + sp] = bv.imag();
    cntOut++;
  } else {
    coh = cplx(0.0, 0.0);
  }
}

// RiccatiSolverConfined.h:228-394 -- both directions for angle k.
// storeGrid = false is the burn-in kernel.
static long long riccatiAngle(Oracle &o, int k, const std::vector<double> &drot,
                              const std::vector<double> &es, bool storeGrid) {
  const int N1 = o.N1, Ne = o.Ne;
  const size_t n1 = o.n1sq();
  const FrameGeometry &g = o.frames[k];
  const double step = o.h / 1.0;
  const int x0 = o.xmin[k], x1 = o.xmax[k], y0 = o.ymin[k], y1 = o.ymax[k];
  const int W = x1 - x0 + 1, H = y1 - y0 + 1;
  if (storeGrid && o.cfg.zero_stale) {
    std::fill(o.gamma.begin(), o.gamma.end(), 0.0);
    std::fill(o.gammat.begin(), o.gammat.end(), 0.0);
  }
  const long long total = 2LL * W * Ne;
#pragma omp parallel for schedule(dynamic, 1)
  for (long long t = 0; t < total; ++t) {
    const bool isGamma = t < (long long)W * Ne;
    const long long r = isGamma ? t : t - (long long)W * Ne;
    const int x = x0 + int(r % W);
    const int e = int(r / W);
    const cplx e0(o.energyRe[e], o.energyIm[e]);
    int cIn = 0, cOut = 0;
    cplx eprev(0, 0), dprev(0, 0), coh(0, 0);
    double *gr = isGamma ? o.gamma.data() : o.gammat.data();
    for (int i = 0; i < H; ++i) {
      const int y = isGamma ? y0 + i : y1 - i;
      const size_t idx = size_t(y) * N1 + x;
      const cplx en = e0 + es[idx];
      const cplx dl = isGamma ? cplx(drot[idx], drot[n1 + idx])
                              : cplx(-drot[idx], drot[n1 + idx]);
      const signed char lab = g.label[idx];
      if (isGamma)
        siteStep<true>(o, step, k, e, x, y, en, eprev, dl, dprev, lab,
                       g.dist.data(), o.idxInG.data(), o.idxOutG.data(),
                       o.bInG.data(), o.nInG, o.bOutG.data(), o.nOutG, cIn,
                       cOut, coh);
      else
        siteStep<false>(o, step, k, e, x, y, en, eprev, dl, dprev, lab,
                        g.dist.data(), o.idxInGt.data(), o.idxOutGt.data(),
                        o.bInGt.data(), o.nInGt, o.bOutGt.data(), o.nOutGt,
                        cIn, cOut, coh);
      if (storeGrid) {
        const size_t gi = size_t(e) * 2 * n1 + idx;
        gr[gi] = coh.real();
        gr[gi + n1] = coh.imag();
      }
      eprev = en;
      dprev = dl;
    }
  }
  // trajectory steps at the sites inside the grain (label > 0)
  long long inside = 0;
  for (int y = y0; y <= y1; ++y)
    for (int x = x0; x <= x1; ++x) inside += g.label[size_t(y) * N1 + x] > 0 ? 1 : 0;
  return 2LL * Ne * inside;
}

// GridFunctions_GPU.h:1121-1189 -- threads per block along x for a 2D
// launch of extent (X, Y).
static int blockDimX(int X, int Y) {
  const int maxT = 128;
  int nx = std::min(static_cast<int>(std::pow(2, static_cast<int>(std::ceil(std::log2(X))))), maxT);
  int ny = std::min(static_cast<int>(std::pow(2, static_cast<int>(std::ceil(std::log2(Y))))), maxT);
  int nz = 1;
  while (nx * ny * nz > maxT) {
    const bool yL = (ny >= nz) && (ny >= nx);
    const bool zL = (nz >= ny) && (nz >= nx);
    if (zL) nz = std::max(nz / 2, 1);
    else if (yL) ny = std::max(ny / 2, 1);
    else nx = std::max(nx / 2, 1);
  }
  return nx;
}

// RiccatiSolverConfined.h:400-468,673-767 -- homogeneous values at exits.
// The launch covers x in [xmin, xmin + ceil(W/bx)*bx) restricted to x < N1.
static void initialConditionAngle(Oracle &o, int k,
                                  const std::vector<double> &drot,
                                  const std::vector<double> &es) {
  const int N1 = o.N1, Ne = o.Ne;
  const size_t n1 = o.n1sq();
  const FrameGeometry &g = o.frames[k];
  const int W = o.xmax[k] - o.xmin[k] + 1;
  const int bx = blockDimX(W, Ne);
  const int xEnd = std::min(N1, o.xmin[k] + ((W + bx - 1) / bx) * bx);
#pragma omp parallel for schedule(static)
  for (int x = o.xmin[k]; x < xEnd; ++x) {
    for (int dir = 0; dir < 2; ++dir) {
      const bool isGamma = dir == 0;
      for (int e = 0; e < Ne; ++e) {
        const cplx e0(o.energyRe[e], o.energyIm[e]);
        int cnt = 0;
        for (int i = 0; i < N1; ++i) {
          const int y = isGamma ? i : N1 - 1 - i;
          const size_t idx = size_t(y) * N1 + x;
          const cplx en = e0 + es[idx];
          const signed char lab = g.label[idx];
          if (lab == (isGamma ? 3 : 2)) {
            const cplx dl(drot[idx], drot[n1 + idx]);
            const int *base = isGamma ? o.idxOutG.data() : o.idxOutGt.data();
            const size_t sp = sparseIdx(base, N1, k, x, cnt, Ne, e);
            cplx ga, gt;
            homogeneous(dl, en, ga, gt);
            const cplx gb = isGamma ? ga : gt;
            double *bo = isGamma ? o.bOutG.data() : o.bOutGt.data();
            const long long nout = isGamma ? o.nOutG : o.nOutGt;
            bo[sp] = gb.real();
            bo[size_t(nout * Ne) + sp] = gb.imag();
            cnt++;
          }
        }
      }
    }
  }
}

// BoundaryConditionSpecular.h:108-240
static void bcInterpolate(const Oracle &o, int ix, int iy, int ie, double sa,
                          double rmx, double rmy, double coordOff, double bdist,
                          int dir, const int *idxOut, const int *numIs,
                          const int *coord, const double *bout, long long nout,
                          double &ore, double &oim, double &ow) {
  const int N1 = o.N1, Ne = o.Ne, Na = o.Na;
  double mx, my;
  rotate(sa, rmx, rmy, mx, my);
  const double fai = fractionalAngleIndex(mx, my, Na);
  const int a0 = wrapToRangeInt(static_cast<int>(fai), 0, Na);
  const int a1 = wrapToRangeInt(static_cast<int>(fai) + 1, 0, Na);
  const int ai[2] = {a0, a1};
  const double pra[2] = {sa - o.systemAngle[a0], sa - o.systemAngle[a1]};
  const double frac = fai - std::floor(fai);
  const double aw[2] = {1.0 - frac, frac};
  const double bpx = static_cast<double>(ix) - coordOff;
  const double bpy = static_cast<double>(iy) -
                     static_cast<double>(sgn(static_cast<double>(dir))) * bdist -
                     coordOff;
  ore = 0.0;
  oim = 0.0;
  ow = 0.0;
  for (int rot = 0; rot < 2; ++rot) {
    double rx, ry;
    rotate(pra[rot], bpx, bpy, rx, ry);
    rx += coordOff;
    ry += coordOff;
    const int pi[2] = {static_cast<int>(rx), static_cast<int>(rx) + 1};
    const double pf = rx - std::floor(rx);
    const double pw[2] = {1.0 - pf, pf};
    for (int pt = 0; pt < 2; ++pt) {
      const int si = ai[rot] * N1 + pi[pt];
      // Boundary points lie well inside the N1 frame (grain margin plus the
      // rotation margin), so si is always a valid column; guard anyway.
      if (pi[pt] < 0 || pi[pt] >= N1) continue;
      const int nis = numIs[si];
      int best = 0;
      double minDy = 1e6;
      for (int i = 0; i < nis; ++i) {
        const int ai2 = i + idxOut[si];
        const double dy = std::fabs(static_cast<double>(coord[ai2]) - ry);
        if (dy < minDy) {
          best = i;
          minDy = dy;
        }
      }
      const long long sp =
          (long long)(idxOut[si] + best) * Ne + ie; // getSparseIndex
      if (sp >= 0) {
        const double w = pw[pt] * aw[rot];
        ore += bout[sp] * w;
        oim += bout[size_t(nout * Ne) + sp] * w;
        ow += w;
      }
    }
  }
}

// BoundaryConditionSpecular.h:245-358,361-486
static double boundaryCondition(Oracle &o) {
  const int N1 = o.N1, Ne = o.Ne, Na = o.Na;
  const std::vector<double> prevG = o.bInG, prevGt = o.bInGt;
  const double coordOff = static_cast<double>(N1 - 1) * 0.5;
  for (int k = 0; k < Na; ++k) {
    const double sa = o.systemAngle[k];
    const double phi = polarAngle(k, Na);
    double mx = std::cos(phi), my = std::sin(phi);
    {
      double tx, ty;
      rotate(-sa, mx, my, tx, ty);
      mx = tx;
      my = ty;
    }
    const FrameGeometry &g = o.frames[k];
    const int xa = o.xmin[k], xb = o.xmax[k], ya = o.ymin[k], yb = o.ymax[k];
    for (int dir = 0; dir < 2; ++dir) {
      const bool isGamma = dir == 0;
      const signed char blabel = isGamma ? 2 : 3;
      const int direction = isGamma ? 1 : -1;
      const int *idxIn = isGamma ? o.idxInG.data() : o.idxInGt.data();
      const int *idxOut = isGamma ? o.idxOutG.data() : o.idxOutGt.data();
      const int *numIs = isGamma ? o.numIsG.data() : o.numIsGt.data();
      const int *coord = isGamma ? o.coordG.data() : o.coordGt.data();
      double *bin = isGamma ? o.bInG.data() : o.bInGt.data();
      const double *bout = isGamma ? o.bOutG.data() : o.bOutGt.data();
      const long long nin = isGamma ? o.nInG : o.nInGt;
      const long long nout = isGamma ? o.nOutG : o.nOutGt;
#pragma omp parallel for schedule(dynamic, 4) collapse(2)
      for (int x = xa; x <= xb; ++x)
        for (int e = 0; e < Ne; ++e) {
          int nEntry = 0;
          const int ny = yb - ya + 1;
          for (int yy = 0; yy < ny; ++yy) {
            const int y = direction > 0 ? (ya + yy) : (yb - yy);
            const size_t idx = size_t(y) * N1 + x;
            if (g.label[idx] != blabel) continue;
            const double bd = std::fabs(g.dist[idx]);
            const double two = 2.0 * (mx * g.nx[idx] + my * g.ny[idx]);
            const double rmx = mx - two * g.nx[idx];
            const double rmy = my - two * g.ny[idx];
            double ir, ii, tw;
            bcInterpolate(o, x, y, e, sa, rmx, rmy, coordOff, bd, direction,
                          idxOut, numIs, coord, bout, nout, ir, ii, tw);
            if (tw < 1e-7) {
              bcInterpolate(o, x, y, e, sa, -mx, -my, coordOff, bd, direction,
                            idxOut, numIs, coord, bout, nout, ir, ii, tw);
            }
            const size_t sp = sparseIdx(idxIn, N1, k, x, nEntry, Ne, e);
            if (tw > 0.0) {
              bin[sp] = ir / tw;
              bin[size_t(nin * Ne) + sp] = ii / tw;
            } else {
              bin[sp] = 0.0;
              bin[size_t(nin * Ne) + sp] = 0.0;
            }
            nEntry++;
          }
        }
    }
  }
  const int norm = o.cfg.norm;
  const double rg = residualArrays(prevG.data(), o.bInG.data(), prevG.size(), norm);
  const double rgt =
      residualArrays(prevGt.data(), o.bInGt.data(), prevGt.size(), norm);
  o.resBoundary = std::max(rg, rgt);
  return o.resBoundary;
}

// IntegrationIteratorOzaki.h:244-300
static void setBoundaryInitialCondition(Oracle &o) {
  std::vector<double> drot, es;
  for (int k = 0; k < o.Na; ++k) {
    rotatedOrderParameter(o, k, drot);
    energyShift(o, k, es);
    initialConditionAngle(o, k, drot, es);
  }
  // RiccatiSolver::computeBoundaryConditions (RiccatiSolver.h:256-277)
  boundaryCondition(o);
  o.boundaryInit = true;
}

// GreensFunctionStatic.h:125-157 + ComputeFreeEnergy.h:106-160
static void greensAngle(const Oracle &o, const std::vector<double> &drot,
                        std::vector<double> &gD, std::vector<double> &gJ,
                        std::vector<double> &gF) {
  const size_t n1 = o.n1sq();
  gD.assign(2 * n1, 0.0);
  gJ.assign(2 * n1, 0.0);
  gF.assign(2 * n1, 0.0);
  const int Ne = o.Ne;
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n1; ++i) {
    cplx sD(0, 0), sJ(0, 0), sF(0, 0);
    const cplx dl(drot[i], drot[n1 + i]);
    for (int e = 0; e < Ne; ++e) {
      const size_t gi = size_t(e) * 2 * n1 + i;
      const cplx ga(o.gamma[gi], o.gamma[gi + n1]);
      const cplx gt(o.gammat[gi], o.gammat[gi + n1]);
      sD += o.residue[e] * greensDelta(ga, gt);
      sJ += o.residue[e] * greensG(ga, gt);
      sF += o.residue[e] * (cmul(std::conj(dl), ga) - cmul(dl, gt));
    }
    gD[i] += sD.real();
    gD[n1 + i] += sD.imag();
    gJ[i] += sJ.real();
    gJ[n1 + i] += sJ.imag();
    gF[i] -= sF.imag();
    gF[n1 + i] += sF.real();
  }
}

// ComputeVectorPotential.h:212-294 (analytic matrix elements)
static double helperAlpha(int a, int b) {
  const double d = std::hypot(static_cast<double>(a), static_cast<double>(b));
  const double f = static_cast<double>(a * b);
  return f * (2.0 * std::log(d) - std::log(2.0));
}
static double helperBeta(int a, int b) {
  const double f = 0.5 * static_cast<double>((2 - a + b) * (a + b) - 2);
  const double ang = static_cast<double>(a) / static_cast<double>(b);
  return f * std::atan(ang);
}
// ComputeVectorPotential.h:246-294: the analytic matrix element, or (the
// reference's inAnalytical = false) the value of 2 ln r at lattice offsets
// with the origin's own value.
static double matrixElement(int i, int j, bool analytical = true) {
  if (!analytical) {
    if (i == 0 && j == 0) return 0.5 * M_PI - std::log(2.0) - 3.0;
    return 2.0 * std::log(std::hypot(static_cast<double>(i), static_cast<double>(j)));
  }
  const int iv[2] = {-i, i}, jv[2] = {-j, j};
  double s = 0.0;
  for (int jj : jv) {
    const int b = 1 + 2 * jj;
    for (int ii : iv) {
      const int a = 1 + 2 * ii;
      s += helperAlpha(a, b) + helperBeta(a, b) + helperBeta(b, a);
    }
  }
  return s * 0.25;
}
// Eigen-decomposition of a symmetric n x n matrix S (row-major, destroyed)
// by the cyclic Jacobi method (Jacobi 1846; rotations in row order, sweeps
// until every off-diagonal element is negligible): eigenvalues in w,
// eigenvector k in column k of V (row-major n x n).
static void jacobiEigen(int n, std::vector<double> &S, std::vector<double> &w,
                        std::vector<double> &V) {
  const size_t N = size_t(n);
  V.assign(N * N, 0.0);
  for (size_t i = 0; i < N; ++i) V[i * N + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (size_t p = 0; p < N; ++p) {
      diag += S[p * N + p] * S[p * N + p];
      for (size_t q = p + 1; q < N; ++q) off += S[p * N + q] * S[p * N + q];
    }
    if (off <= 1e-32 * diag) break;
    for (size_t p = 0; p + 1 < N; ++p)
      for (size_t q = p + 1; q < N; ++q) {
        const double apq = S[p * N + q];
        if (apq == 0.0) continue;
        const double app = S[p * N + p], aqq = S[q * N + q];
        // the rotation angle that zeroes S[p][q]: tan(2 phi) = 2 apq / (aqq - app)
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) /
                         (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
        for (size_t k = 0; k < N; ++k) { // columns p and q
          const double skp = S[k * N + p], skq = S[k * N + q];
          S[k * N + p] = c * skp - sn * skq;
          S[k * N + q] = sn * skp + c * skq;
        }
        for (size_t k = 0; k < N; ++k) { // rows p and q
          const double spk = S[p * N + k], sqk = S[q * N + k];
          S[p * N + k] = c * spk - sn * sqk;
          S[q * N + k] = sn * spk + c * sqk;
        }
        S[p * N + q] = 0.0;
        S[q * N + p] = 0.0;
        for (size_t k = 0; k < N; ++k) {
          const double vkp = V[k * N + p], vkq = V[k * N + q];
          V[k * N + p] = c * vkp - sn * vkq;
          V[k * N + q] = sn * vkp + c * vkq;
        }
      }
  }
  w.resize(N);
  for (size_t i = 0; i < N; ++i) w[i] = S[i * N + i];
}
// ComputeVectorPotential.h:212-244,296-417 -- the 2D filter: the full
// kernel (maxError = 0) or C plus the eigen-expansion of Q truncated by
// magnitude at relative Frobenius error maxError (the sum of the
// reference's separable filters).  Returns the number of rank-one filters.
static int makeFilter2D(int N, double h, std::vector<double> &filter,
                        double maxError = 0.0, bool analytical = true) {
  std::vector<double> Q(size_t(N) * N);
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i) Q[size_t(j) * N + i] = matrixElement(i, j, analytical);
  // Q(i,j) stored at [j*N + i] (column-major like the reference); symmetrize
  std::vector<double> S(size_t(N) * N);
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i)
      S[size_t(j) * N + i] = 0.5 * (Q[size_t(j) * N + i] + Q[size_t(i) * N + j]);
  const double scale = 1.0 / (4.0 * M_PI);
  for (double &v : S) v *= scale;
  double C = (analytical ? (2.0 * std::log(h) - std::log(2.0) - 3.0) : 2.0 * std::log(h)) * scale;
  // arma::mean(arma::mean(Q)): mean of column means.
  double mean = 0.0;
  for (int j = 0; j < N; ++j) {
    double cm = 0.0;
    for (int i = 0; i < N; ++i) cm += S[size_t(j) * N + i];
    mean += cm / N;
  }
  mean /= N;
  for (double &v : S) v -= mean;
  C += mean;
  int filters = 0;
  if (maxError > 0.0) {
    std::vector<double> w, V;
    jacobiEigen(N, S, w, V);
    // arma::sort_index(abs(eig), "descend"), added until the error is met
    std::vector<int> order(static_cast<size_t>(N));
    for (int i = 0; i < N; ++i) order[size_t(i)] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      return std::fabs(w[size_t(a)]) > std::fabs(w[size_t(b)]);
    });
    double norm2 = 0.0;
    for (double v : w) norm2 += v * v;
    std::vector<int> kept;
    double partial = 0.0;
    for (int k : order) {
      kept.push_back(k);
      partial += w[size_t(k)] * w[size_t(k)];
      if (std::sqrt(std::max(1.0 - partial / norm2, 0.0)) < maxError) break;
    }
    const bool withC = std::fabs(C) > 0.0;
    std::fill(S.begin(), S.end(), withC ? C : 0.0);
    for (int k : kept)
      for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i)
          S[size_t(j) * N + i] += w[size_t(k)] * V[size_t(i) * N + size_t(k)] *
                                  V[size_t(j) * N + size_t(k)];
    filters = static_cast<int>(kept.size()) + (withC ? 1 : 0);
  } else {
    for (double &v : S) v += C;
  }
  filter.resize(size_t(2 * N - 1) * (2 * N - 1));
  size_t p = 0;
  for (int j = -N + 1; j < N; ++j)
    for (int i = -N + 1; i < N; ++i) {
      const int idx = N * std::abs(j) + std::abs(i);
      filter[p++] = S[idx];
    }
  return filters;
}
// ComputeVectorPotential.h:519-563,575-616,647-745
static void solveVectorPotential(int N, double h, int numGrain, bool enforce,
                                 const std::vector<double> &filter,
                                 const double *jin, double *out) {
  const size_t n = size_t(N) * N;
  std::vector<double> src(jin, jin + 2 * n);
  if (enforce) {
    for (int d = 0; d < 2; ++d) {
      double s = 0.0;
      for (size_t i = 0; i < n; ++i) s += jin[d * n + i];
      const double mean = s / static_cast<double>(numGrain);
      for (size_t i = 0; i < n; ++i) src[d * n + i] = jin[d * n + i] - mean;
    }
  }
  const int W = 2 * N - 1;
#pragma omp parallel for collapse(2) schedule(static)
  for (int d = 0; d < 2; ++d)
    for (int y = 0; y < N; ++y)
      for (int x = 0; x < N; ++x) {
        const int ofx = N - 1 - x, ofy = N - 1 - y;
        double sum = 0.0;
        for (int j = 0; j < N; ++j) {
          double ps = 0.0;
          const double *frow = &filter[size_t(W) * (j + ofy) + ofx];
          const double *srow = &src[d * n + size_t(N) * j];
#pragma omp simd reduction(+ : ps)
          for (int i = 0; i < N; ++i) ps += srow[i] * frow[i];
          sum += ps;
        }
        out[d * n + size_t(N) * y + x] = sum;
      }
  const double f = -1.0 * (h * h) * 1.0;
  for (size_t i = 0; i < 2 * n; ++i) out[i] *= f;
}

// curl_kernel (GridFunctions_GPU.h:326-353)
static double sumCurlSquare(const Oracle &o) {
  const int N = o.N0;
  const size_t n = o.n0sq();
  const double inv = 1.0 / o.h;
  const double *Ax = o.A.data(), *Ay = o.A.data() + n;
  double s = 0.0;
  for (int y = 0; y < N; ++y)
    for (int x = 0; x < N; ++x) {
      const int xn = std::min(x + 1, N - 1), xp = std::max(x - 1, 0);
      const int yn = std::min(y + 1, N - 1), yp = std::max(y - 1, 0);
      const double dAydx =
          (Ay[size_t(y) * N + xn] - Ay[size_t(y) * N + xp]) / double(xn - xp);
      const double dAxdy =
          (Ax[size_t(yn) * N + x] - Ax[size_t(yp) * N + x]) / double(yn - yp);
      const double b = (dAydx - dAxdy) * inv;
      s += b * b;
    }
  return s;
}

static void zeroOutside(const Oracle &o, double *plane) {
  for (size_t i = 0; i < o.n0sq(); ++i)
    if (o.domainRef[i] == 0) plane[i] = 0.0;
}

// IntegrationIteratorOzaki.h:303-404 + compute objects.
// One angle of IntegrationIteratorOzaki::compute (IntegrationIteratorOzaki.h:
// 300-420): rotate the order parameter, energy shift, Riccati sweep, Green
// functions, and their integration into the angle sums.  Returns the number
// of trajectory steps taken.
static long long integrateAngle(Oracle &o, int k, std::vector<double> &drot,
                                std::vector<double> &es, std::vector<double> &gD,
                                std::vector<double> &gJ, std::vector<double> &gF,
                                std::vector<double> &tmp) {
  const size_t n0 = o.n0sq(), n1 = o.n1sq();
  rotatedOrderParameter(o, k, drot);
  energyShift(o, k, es);
  const long long steps = riccatiAngle(o, k, drot, es, true);
  greensAngle(o, drot, gD, gJ, gF);
  const signed char *lab = o.frames[k].label.data();
  const double sa = o.systemAngle[k];
  const float phi = static_cast<float>(sa);
  const double wk = angleWeight(k, o.Na) / static_cast<double>(o.Na);
  const double pphi = polarAngle(k, o.Na);
  const double vx = std::cos(pphi), vy = std::sin(pphi);
  // ComputeOrderParameter::integrateGreensAngle (ComputeOrderParameter.h:174-210)
  extendValues(lab, o.N1, gD.data());
  extendValues(lab, o.N1, gD.data() + n1);
  for (int c = 0; c < o.Nc; ++c) {
    std::fill(tmp.begin(), tmp.end(), 0.0);
    insertRotated(tmp.data(), o.N0, o.N0, gD.data(), o.N1, o.N1, phi, false);
    insertRotated(tmp.data() + n0, o.N0, o.N0, gD.data() + n1, o.N1, o.N1,
                  phi, false);
    const double b = basisMomentum(o, c, k);
    const double bi = -0.0; // conj of a real basis value
    double *acc = o.deltaNew[c].data();
    for (size_t i = 0; i < n0; ++i) {
      const double re = tmp[i], im = tmp[n0 + i];
      const double nre = re * b - im * bi, nim = re * bi + im * b;
      acc[i] += nre * wk;
      acc[n0 + i] += nim * wk;
    }
  }
  if (o.cfg.compute_current) {
    // ComputeCurrent::integrateGreensAngle (ComputeCurrent.h:142-173)
    extendValues(lab, o.N1, gJ.data());
    extendValues(lab, o.N1, gJ.data() + n1);
    for (int f = 0; f < 2; ++f) {
      const double fac = (f == 0 ? vx : vy) * wk;
      std::vector<double> t2(gJ);
      for (double &v : t2) v *= fac;
      insertRotated(o.currentAcc.data() + f * 2 * n0, o.N0, o.N0, t2.data(),
                    o.N1, o.N1, phi, true);
      insertRotated(o.currentAcc.data() + f * 2 * n0 + n0, o.N0, o.N0,
                    t2.data() + n1, o.N1, o.N1, phi, true);
    }
  }
  if (o.cfg.compute_free_energy) {
    // ComputeFreeEnergy::integrateGreensAngle (ComputeFreeEnergy.h:260-279)
    extendValues(lab, o.N1, gF.data());
    extendValues(lab, o.N1, gF.data() + n1);
    for (double &v : gF) v *= wk;
    insertRotated(o.feGrid.data(), o.N0, o.N0, gF.data(), o.N1, o.N1, phi,
                  true);
    insertRotated(o.feGrid.data() + n0, o.N0, o.N0, gF.data() + n1, o.N1,
                  o.N1, phi, true);
  }
  return steps;
}

static void integrate(Oracle &o, bool burnin) {
  if (!o.boundaryInit) setBoundaryInitialCondition(o);
  const size_t n0 = o.n0sq();
  std::vector<double> drot, es, gD, gJ, gF, tmp(2 * n0);
  if (burnin) {
    for (int k = 0; k < o.Na; ++k) {
      rotatedOrderParameter(o, k, drot);
      energyShift(o, k, es);
      riccatiAngle(o, k, drot, es, false);
    }
  } else {
    for (int c = 0; c < o.Nc; ++c)
      std::fill(o.deltaNew[c].begin(), o.deltaNew[c].end(), 0.0);
    std::fill(o.currentAcc.begin(), o.currentAcc.end(), 0.0);
    std::fill(o.feGrid.begin(), o.feGrid.end(), 0.0);
    for (int k = 0; k < o.Na; ++k) integrateAngle(o, k, drot, es, gD, gJ, gF, tmp);
    // ComputeOrderParameter::computePropertyFromGreens (:212-262)
    double res = 0.0;
    for (int c = 0; c < o.Nc; ++c) {
      for (double &v : o.deltaNew[c]) v *= o.coupling[c];
      zeroOutside(o, o.deltaNew[c].data());
      zeroOutside(o, o.deltaNew[c].data() + n0);
      res = std::max(res, residualArrays(o.delta[c].data(), o.deltaNew[c].data(),
                                         2 * n0, o.cfg.norm));
    }
    o.resOP = res;
    o.convOP = res < o.cfg.convergence_criterion;
    if (o.cfg.compute_current) {
      // ComputeCurrent::computePropertyFromGreens (ComputeCurrent.h:175-222)
      const std::vector<double> old = o.current;
      const double fac = static_cast<double>(2 * sgn(o.cfg.charge_sign)) * o.T;
      for (int f = 0; f < 2; ++f) {
        for (size_t i = 0; i < n0; ++i)
          o.current[f * n0 + i] = o.currentAcc[f * 2 * n0 + i] * fac;
        zeroOutside(o, o.current.data() + f * n0);
      }
      const double actual =
          residualArrays(old.data(), o.current.data(), 2 * n0, o.cfg.norm);
      o.resJ = o.firstCurrent ? o.cfg.convergence_criterion * 0.1 : actual;
      o.convJ = o.resJ < o.cfg.convergence_criterion;
      o.firstCurrent = false;
    }
    if (o.cfg.compute_free_energy) {
      // ComputeFreeEnergy::computePropertyFromGreens (:281-414)
      double sumInv = 0.0;
      for (int e = 0; e < o.Ne; ++e) sumInv += o.residue[e] / o.energyIm[e];
      sumInv *= o.T;
      zeroOutside(o, o.feGrid.data());
      zeroOutside(o, o.feGrid.data() + n0);
      double sre = 0.0;
      for (size_t i = 0; i < n0; ++i) sre += o.feGrid[i];
      double nonMag = o.T * sre;
      for (int c = 0; c < o.Nc; ++c) {
        const double factor = std::log(o.T / o.tc[c]) + sumInv;
        double s2 = 0.0;
        for (size_t i = 0; i < n0; ++i) {
          if (o.domainRef[i] == 0) continue;
          const double re = o.delta[c][i], im = o.delta[c][n0 + i];
          s2 += re * re + im * im;
        }
        nonMag += s2 * factor;
      }
      double mag = 0.0;
      if (o.lambda > 0.0) mag = 0.5 / (o.lambda * o.lambda) * sumCurlSquare(o);
      const double area = o.h * o.h;
      const double perArea = (nonMag + mag) / static_cast<double>(o.numGrain);
      const double actual = residualScalar(o.fePerArea, perArea);
      o.resFE = o.firstFE ? o.cfg.convergence_criterion * 0.1 : actual;
      o.convFE = o.resFE < o.cfg.convergence_criterion;
      o.feTotal = (nonMag + mag) * area;
      o.fePerArea = perArea;
      o.feMag = mag * area;
      o.feMagPerArea = mag / static_cast<double>(o.numGrain);
      o.firstFE = false;
    }
  }
  boundaryCondition(o);
}

// ---------------------------------------------------------------------------
// Discrete adjoint of the free energy (no counterpart in the reference).
// Omega(Delta) is the free energy one iteration reports (ComputeFreeEnergy.h:
// 281-414 on the Green functions of the sweep), with the vector potential and
// the boundary in-values held fixed.  Reverse mode, written from the
// mathematics of each forward stage: the back-rotation and extension of the
// free-energy density, the Green sums, the two Riccati recurrences, and the
// rotation and extension of the order parameter.  For a complex variable z
// the adjoint is zbar = dOmega/dRe z + i dOmega/dIm z; for y = f(x) with f
// holomorphic, xbar = conj(f'(x)) ybar.
// ---------------------------------------------------------------------------

// Adjoint of riccatiStep(delta, energy, y0, h).  Its result y is the root of
// F(y) = a (y + y0)^2 + (se - 1) y + (se + 1) y0 + s delta = 0 with s = i h,
// a = s conj(delta) / 4 and se = s energy, so by implicit differentiation
// dy = -(F_y0 dy0 + F_a da + F_sd d(s delta)) / F_y with
// F_y = 2 a (y + y0) + se - 1, F_y0 = 2 a (y + y0) + se + 1, F_a = (y + y0)^2.
// F_se = y + y0, so the energy's adjoint is conj(q s (y + y0)) ybar.
// Returns y0bar and adds the adjoints of delta and of the energy to dbar and
// ebar.
static cplx riccatiStepAdjoint(cplx delta, cplx energy, cplx y0, cplx y, double h,
                               cplx ybar, cplx &dbar, cplx &ebar) {
  const cplx s(0.0, h);
  const cplx a = 0.25 * cmul(s, std::conj(delta));
  const cplx se = cmul(s, energy);
  const cplx w = y + y0;
  const cplx aw2 = 2.0 * cmul(a, w);
  const cplx q = -cdiv(cplx(1.0, 0.0), aw2 + se - 1.0);
  const cplx abar = cmul(std::conj(cmul(q, cmul(w, w))), ybar);
  const cplx sdbar = cmul(std::conj(q), ybar);
  // a = (s / 4) conj(delta): deltabar += conj(abar) s / 4; s delta: conj(s) sdbar
  dbar += cmul(std::conj(abar), 0.25 * s) + cmul(std::conj(s), sdbar);
  ebar += cmul(std::conj(cmul(q, cmul(s, w))), ybar);
  return cmul(std::conj(cmul(q, aw2 + se + 1.0)), ybar);
}

// Adjoint of insertRotated: the adjoint of every written destination site is
// scattered to its four source taps with the bilinear weights.
static void insertRotatedAdjoint(const double *dstbar, int Nx, int Ny, double *srcbar, int Nxo,
                                 int Nyo, float phi) {
  const double xt = (Nxo - Nx) / 2.0, yt = (Nyo - Ny) / 2.0;
  const double sx = (Nxo - 1) / 2.0, sy = (Nyo - 1) / 2.0;
  const double c = static_cast<double>(std::cos(-phi));
  const double s = static_cast<double>(std::sin(-phi));
  for (int y = 0; y < Ny; ++y)
    for (int x = 0; x < Nx; ++x) {
      const double v = dstbar[size_t(y) * Nx + x];
      if (v == 0.0) continue;
      const double X = x + xt - sx;
      const double Y = y + yt - sy;
      const double Xr = X * c - Y * s + sx;
      const double Yr = X * s + Y * c + sy;
      const int Xi = static_cast<int>(Xr), Yi = static_cast<int>(Yr);
      if (Xi >= 0 && Xi + 1 < Nxo && Yi >= 0 && Yi + 1 < Nyo) {
        const double fx = Xr - (double)Xi, fy = Yr - (double)Yi;
        srcbar[size_t(Yi) * Nxo + Xi] += (1.0 - fy) * (1.0 - fx) * v;
        srcbar[size_t(Yi) * Nxo + Xi + 1] += (1.0 - fy) * fx * v;
        srcbar[size_t(Yi + 1) * Nxo + Xi] += fy * (1.0 - fx) * v;
        srcbar[size_t(Yi + 1) * Nxo + Xi + 1] += fy * fx * v;
      }
    }
}

// Adjoint of extendValues: a label-0 site that took the mean of its nearest
// inside sites hands its adjoint to them in equal parts and keeps none.
static void extendValuesAdjoint(const signed char *labels, int N, double *bar) {
  const int ext = 2;
  for (int y = 0; y < N; ++y)
    for (int x = 0; x < N; ++x) {
      if (labels[size_t(y) * N + x] != 0) continue;
      const int x0 = std::max(x - ext, 0), x1 = std::min(x + ext, N - 1);
      const int y0 = std::max(y - ext, 0), y1 = std::min(y + ext, N - 1);
      int best = 1000, np = 0;
      for (int ky = y0; ky <= y1; ++ky)
        for (int kx = x0; kx <= x1; ++kx)
          if (labels[size_t(ky) * N + kx] > 0) {
            const int dx = x - kx, dy = y - ky, d2 = dx * dx + dy * dy;
            if (d2 < best) {
              best = d2;
              np = 1;
            } else if (d2 == best) {
              np++;
            }
          }
      if (np == 0) continue;
      const double v = bar[size_t(y) * N + x] / static_cast<double>(np);
      bar[size_t(y) * N + x] = 0.0;
      for (int ky = y0; ky <= y1; ++ky)
        for (int kx = x0; kx <= x1; ++kx)
          if (labels[size_t(ky) * N + kx] > 0) {
            const int dx = x - kx, dy = y - ky;
            if (dx * dx + dy * dy == best) bar[size_t(ky) * N + kx] += v;
          }
    }
}

// Reverse sweep of one column (x) and energy (e) of frame k for one
// coherence function: grid holds the forward values at every site of the
// column, rho the adjoint seed of the free-energy density per frame site.
// Adds the adjoint of the frame's order parameter to dbar (planar complex)
// and that of its Doppler shift (a real shift of every energy) to esbar.
// With loc, the local adjoint of the coherence function at every site and
// energy (loc[e n1 + site]) replaces the free-energy density's.
template <bool GAMMA>
static void reverseColumn(const Oracle &o, int k, int x, int e, const std::vector<double> &drot,
                          const std::vector<double> &es, const std::vector<double> &rho,
                          std::vector<double> &dbar, std::vector<double> &esbar,
                          const std::vector<cplx> *loc = nullptr) {
  const int N1 = o.N1, Ne = o.Ne;
  const size_t n1 = o.n1sq();
  const FrameGeometry &g = o.frames[k];
  const int ENTRY = GAMMA ? 2 : 3, EXIT = GAMMA ? 3 : 2;
  const int ya = o.ymin[k], yb = o.ymax[k], H = yb - ya + 1;
  const double step = o.h;
  const double *grid = (GAMMA ? o.gamma.data() : o.gammat.data()) + size_t(e) * 2 * n1;
  const int *idxIn = GAMMA ? o.idxInG.data() : o.idxInGt.data();
  const double *bin = GAMMA ? o.bInG.data() : o.bInGt.data();
  const long long nin = GAMMA ? o.nInG : o.nInGt;
  const cplx e0(o.energyRe[e], o.energyIm[e]);
  auto siteOf = [&](int i) { return size_t(GAMMA ? ya + i : yb - i) * N1 + x; };
  auto dlOf = [&](size_t idx) {
    return GAMMA ? cplx(drot[idx], drot[n1 + idx]) : cplx(-drot[idx], drot[n1 + idx]);
  };
  // entry counts in forward order (the sparse index of each entry)
  std::vector<int> entryCount(size_t(H), 0);
  for (int i = 0, cnt = 0; i < H; ++i) {
    entryCount[size_t(i)] = cnt;
    if (g.label[siteOf(i)] == ENTRY) ++cnt;
  }
  // the adjoint of the argument dl of a step, returned to the true order
  // parameter D (dl = D for gamma, dl = -conj(D) for gamma~)
  auto addDbar = [&](size_t idx, cplx b) {
    const cplx d = GAMMA ? b : -std::conj(b);
    dbar[idx] += d.real();
    dbar[n1 + idx] += d.imag();
  };
  cplx carry(0.0, 0.0);
  for (int i = H - 1; i >= 0; --i) {
    const size_t idx = siteOf(i);
    const cplx D(drot[idx], drot[n1 + idx]);
    // dOmega/dgamma at the site through the free-energy density
    // gF = sum_e r_e [D.re (gt.im - g.im) + D.im (g.re + gt.re)]
    const double rr = loc ? 0.0 : rho[idx] * o.residue[e];
    const cplx local = loc ? (*loc)[size_t(e) * n1 + idx]
                           : (GAMMA ? cplx(rr * D.imag(), -rr * D.real())
                                    : cplx(rr * D.imag(), rr * D.real()));
    const cplx ybar = carry + local;
    const signed char lab = g.label[idx];
    const cplx y(grid[idx], grid[n1 + idx]);
    const cplx en = e0 + es[idx];
    if (lab == 1 || lab == EXIT) {
      cplx y0(0.0, 0.0), dprev(0.0, 0.0), eprev(0.0, 0.0);
      size_t pidx = 0;
      if (i > 0) {
        pidx = siteOf(i - 1);
        y0 = cplx(grid[pidx], grid[n1 + pidx]);
        dprev = dlOf(pidx);
        eprev = e0 + es[pidx];
      }
      cplx db(0.0, 0.0), eb(0.0, 0.0);
      carry = riccatiStepAdjoint((dprev + dlOf(idx)) * 0.5, (eprev + en) * 0.5, y0, y, step, ybar,
                                 db, eb);
      addDbar(idx, 0.5 * db);
      esbar[idx] += 0.5 * eb.real();
      if (i > 0) {
        addDbar(pidx, 0.5 * db);
        esbar[pidx] += 0.5 * eb.real();
      }
    } else if (lab == ENTRY) {
      const size_t sp = sparseIdx(idxIn, N1, k, x, entryCount[size_t(i)], Ne, e);
      const cplx bc(bin[sp], bin[size_t(nin * Ne) + sp]);
      const double d = std::fabs(g.dist[idx]);
      cplx db(0.0, 0.0), eb(0.0, 0.0);
      (void)riccatiStepAdjoint(dlOf(idx), en, bc, y, d * step, ybar, db, eb);
      addDbar(idx, db);
      esbar[idx] += eb.real();
      carry = cplx(0.0, 0.0);
    } else {
      carry = cplx(0.0, 0.0);
    }
  }
}

// Omega and dOmega/dRe Delta_c, dOmega/dIm Delta_c (grad: component c at
// [c * 2 N0^2, re plane then im plane]), and with gradA (2 N0^2: x plane,
// y plane) dOmega/dA of the induced vector potential in the units the
// config's setters take (the state stores lambda^2 A).  Leaves the state as
// it was (the out-values the sweeps write are restored).
static double freeEnergyGradient(Oracle &o, std::vector<double> &grad,
                                 std::vector<double> *gradA = nullptr) {
  if (!o.boundaryInit) setBoundaryInitialCondition(o);
  const std::vector<double> keepG = o.bOutG, keepGt = o.bOutGt;
  const size_t n0 = o.n0sq(), n1 = o.n1sq();
  const int N1 = o.N1, Ne = o.Ne;
  const double area = o.h * o.h;
  const bool withA = gradA != nullptr && o.lambda > 0.0;
  grad.assign(size_t(o.Nc) * 2 * n0, 0.0);
  if (gradA) gradA->assign(2 * n0, 0.0);
  std::vector<double> drot, es, gD, gJ, gF, fe(n0, 0.0), rho0(n0), rho(n1), dbar(2 * n1),
      sumbar(2 * n0), esbar(n1), arbar(n1);
  for (int k = 0; k < o.Na; ++k) {
    rotatedOrderParameter(o, k, drot);
    energyShift(o, k, es);
    riccatiAngle(o, k, drot, es, true);
    greensAngle(o, drot, gD, gJ, gF);
    const signed char *lab = o.frames[k].label.data();
    const float phi = static_cast<float>(o.systemAngle[k]);
    const double wk = angleWeight(k, o.Na) / static_cast<double>(o.Na);
    // forward: the free-energy density of this angle on the reference grid
    extendValues(lab, N1, gF.data());
    for (size_t i = 0; i < n1; ++i) gF[i] *= wk;
    insertRotated(fe.data(), o.N0, o.N0, gF.data(), N1, N1, phi, true);
    // adjoint seed: Omega = area T sum over the inside sites of fe
    for (size_t i = 0; i < n0; ++i) rho0[i] = o.domainRef[i] != 0 ? area * o.T * wk : 0.0;
    std::fill(rho.begin(), rho.end(), 0.0);
    insertRotatedAdjoint(rho0.data(), o.N0, o.N0, rho.data(), N1, N1, phi);
    extendValuesAdjoint(lab, N1, rho.data());
    // the Green sums' own dependence on D, then the two recurrences
    std::fill(dbar.begin(), dbar.end(), 0.0);
    std::fill(esbar.begin(), esbar.end(), 0.0);
    for (size_t i = 0; i < n1; ++i) {
      if (rho[i] == 0.0) continue;
      double su = 0.0, sw = 0.0;
      for (int e = 0; e < Ne; ++e) {
        const size_t gi = size_t(e) * 2 * n1 + i;
        su += o.residue[e] * (o.gammat[gi + n1] - o.gamma[gi + n1]);
        sw += o.residue[e] * (o.gamma[gi] + o.gammat[gi]);
      }
      dbar[i] += rho[i] * su;
      dbar[n1 + i] += rho[i] * sw;
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int x = o.xmin[k]; x <= o.xmax[k]; ++x)
      for (int e = 0; e < Ne; ++e) {
        reverseColumn<true>(o, k, x, e, drot, es, rho, dbar, esbar);
        reverseColumn<false>(o, k, x, e, drot, es, rho, dbar, esbar);
      }
    if (withA) {
      // es = q (v . (A_rot + A_ext)), A_rot the stored lambda^2 A rotated
      // into the frame and divided by lambda^2 (energyShift): the adjoint
      // of A is the rotation's transpose of q v_f esbar
      const double pa = polarAngle(k, o.Na);
      const double vf[2] = {std::cos(pa), std::sin(pa)};
      const int q = sgn(o.cfg.charge_sign);
      for (int f = 0; f < 2; ++f) {
        for (size_t i = 0; i < n1; ++i) arbar[i] = static_cast<double>(q) * vf[f] * esbar[i];
        insertRotatedAdjoint(arbar.data(), N1, N1, gradA->data() + size_t(f) * n0, o.N0, o.N0,
                             static_cast<float>(-o.systemAngle[k]));
      }
    }
    // the rotation and extension of the total order parameter
    std::fill(sumbar.begin(), sumbar.end(), 0.0);
    const int kk = wrapToRangeInt(-k, 0, o.Na);
    const float phiD = static_cast<float>(o.systemAngle[kk]);
    insertRotatedAdjoint(dbar.data(), N1, N1, sumbar.data(), o.N0, o.N0, phiD);
    insertRotatedAdjoint(dbar.data() + n1, N1, N1, sumbar.data() + n0, o.N0, o.N0, phiD);
    extendValuesAdjoint(o.domainRef.data(), o.N0, sumbar.data());
    extendValuesAdjoint(o.domainRef.data(), o.N0, sumbar.data() + n0);
    for (int c = 0; c < o.Nc; ++c) {
      const double b = basisMomentum(o, c, k);
      double *gc = grad.data() + size_t(c) * 2 * n0;
      for (size_t i = 0; i < 2 * n0; ++i) gc[i] += b * sumbar[i];
    }
  }
  // the explicit terms sum_c |Delta_c|^2 (ln(T / Tc_c) + T sum_e r_e / w_e)
  double sumInv = 0.0;
  for (int e = 0; e < o.Ne; ++e) sumInv += o.residue[e] / o.energyIm[e];
  sumInv *= o.T;
  double sre = 0.0;
  for (size_t i = 0; i < n0; ++i)
    if (o.domainRef[i] != 0) sre += fe[i];
  double nonMag = o.T * sre;
  for (int c = 0; c < o.Nc; ++c) {
    const double factor = std::log(o.T / o.tc[c]) + sumInv;
    double s2 = 0.0;
    double *gc = grad.data() + size_t(c) * 2 * n0;
    for (size_t i = 0; i < n0; ++i) {
      if (o.domainRef[i] == 0) continue;
      const double re = o.delta[c][i], im = o.delta[c][n0 + i];
      s2 += re * re + im * im;
      gc[i] += 2.0 * area * factor * re;
      gc[n0 + i] += 2.0 * area * factor * im;
    }
    nonMag += s2 * factor;
  }
  double mag = 0.0;
  if (o.lambda > 0.0) mag = 0.5 / (o.lambda * o.lambda) * sumCurlSquare(o);
  if (withA) {
    // (area / 2 lambda^2) sum b^2 with b the curl of the stored lambda^2 A:
    // per unit of A, area sum_t b_t db_t/dA (sumCurlSquare's stencil)
    const int N = o.N0;
    const double inv = 1.0 / o.h;
    const double *Ax = o.A.data(), *Ay = o.A.data() + n0;
    double *gx = gradA->data(), *gy = gradA->data() + n0;
    for (int y = 0; y < N; ++y)
      for (int x = 0; x < N; ++x) {
        const int xn = std::min(x + 1, N - 1), xp = std::max(x - 1, 0);
        const int yn = std::min(y + 1, N - 1), yp = std::max(y - 1, 0);
        const double cx = inv / double(xn - xp), cy = inv / double(yn - yp);
        const double b = (Ay[size_t(y) * N + xn] - Ay[size_t(y) * N + xp]) * cx -
                         (Ax[size_t(yn) * N + x] - Ax[size_t(yp) * N + x]) * cy;
        gy[size_t(y) * N + xn] += area * b * cx;
        gy[size_t(y) * N + xp] -= area * b * cx;
        gx[size_t(yn) * N + x] -= area * b * cy;
        gx[size_t(yp) * N + x] += area * b * cy;
      }
  }
  o.bOutG = keepG;
  o.bOutGt = keepGt;
  return (nonMag + mag) * area;
}

// The self-consistency residual R = 1/2 sum_inside sum_c |Delta'_c - Delta_c|^2
// + 1/2 sum |A' - A|^2 of one plain iteration u -> G(u) (the gap equation of
// ComputeOrderParameter.h:174-262, the current of ComputeCurrent.h:142-222
// and its vector potential, ComputeVectorPotential.h:647-745), the boundary
// in-values held fixed, and its gradient (J - I)^T (G(u) - u) in Delta
// (grad as in freeEnergyGradient) and in A (gradA, per unit of A as the
// setters take it; zero without a penetration depth).  Reverse mode of the
// same stages as freeEnergyGradient with other seeds: the transposes of the
// Poisson solve, of the current's and the gap equation's back-rotation and
// extension, the local derivatives of the Green functions (LocalGreens.h:
// 32-70) at every site and energy, then the same reverse sweeps.
static double residualGradient(Oracle &o, std::vector<double> &grad, std::vector<double> &gradA) {
  if (!o.boundaryInit) setBoundaryInitialCondition(o);
  const std::vector<double> keepG = o.bOutG, keepGt = o.bOutGt;
  const size_t n0 = o.n0sq(), n1 = o.n1sq();
  const int N1 = o.N1, Ne = o.Ne, Nc = o.Nc;
  const bool doA = o.cfg.compute_current && o.lambda > 0.0;
  const double l2 = doA ? o.lambda * o.lambda : 1.0;
  grad.assign(size_t(Nc) * 2 * n0, 0.0);
  gradA.assign(2 * n0, 0.0);
  std::vector<double> drot, es, gD, gJ, gF, tmp(2 * n0);
  // forward: G(u) from one sweep of every angle
  std::vector<std::vector<double>> dnew(size_t(Nc), std::vector<double>(2 * n0, 0.0));
  std::vector<double> jacc(4 * n0, 0.0), j(2 * n0, 0.0), anew(2 * n0, 0.0);
  const double jfac = static_cast<double>(2 * sgn(o.cfg.charge_sign)) * o.T;
  for (int k = 0; k < o.Na; ++k) {
    rotatedOrderParameter(o, k, drot);
    energyShift(o, k, es);
    riccatiAngle(o, k, drot, es, true);
    greensAngle(o, drot, gD, gJ, gF);
    const signed char *lab = o.frames[k].label.data();
    const float phi = static_cast<float>(o.systemAngle[k]);
    const double wk = angleWeight(k, o.Na) / static_cast<double>(o.Na);
    const double pa = polarAngle(k, o.Na);
    extendValues(lab, N1, gD.data());
    extendValues(lab, N1, gD.data() + n1);
    std::fill(tmp.begin(), tmp.end(), 0.0);
    insertRotated(tmp.data(), o.N0, o.N0, gD.data(), N1, N1, phi, false);
    insertRotated(tmp.data() + n0, o.N0, o.N0, gD.data() + n1, N1, N1, phi, false);
    for (int c = 0; c < Nc; ++c) {
      const double b = basisMomentum(o, c, k);
      for (size_t i = 0; i < 2 * n0; ++i) dnew[size_t(c)][i] += tmp[i] * b * wk;
    }
    if (doA) {
      extendValues(lab, N1, gJ.data());
      for (int f = 0; f < 2; ++f) {
        const double fac = (f == 0 ? std::cos(pa) : std::sin(pa)) * wk;
        std::vector<double> t2(gJ.begin(), gJ.begin() + n1);
        for (double &v : t2) v *= fac;
        insertRotated(jacc.data() + size_t(f) * n0, o.N0, o.N0, t2.data(), N1, N1, phi, true);
      }
    }
  }
  std::vector<std::vector<double>> rD(size_t(Nc), std::vector<double>(2 * n0, 0.0));
  double R = 0.0;
  for (int c = 0; c < Nc; ++c)
    for (size_t i = 0; i < n0; ++i) {
      if (o.domainRef[i] == 0) continue;
      for (int q = 0; q < 2; ++q) {
        const double v = dnew[size_t(c)][q * n0 + i] * o.coupling[c] - o.delta[c][q * n0 + i];
        rD[size_t(c)][q * n0 + i] = v;
        R += 0.5 * v * v;
      }
    }
  std::vector<double> rA(2 * n0, 0.0), jbar(2 * n0, 0.0);
  if (doA) {
    for (int f = 0; f < 2; ++f)
      for (size_t i = 0; i < n0; ++i)
        j[f * n0 + i] = o.domainRef[i] != 0 ? jacc[f * n0 + i] * jfac : 0.0;
    if (o.filter2D.empty())
      makeFilter2D(o.N0, o.h, o.filter2D, std::max(o.cfg.vector_potential_error, 0.0));
    solveVectorPotential(o.N0, o.h, o.numGrain, true, o.filter2D, j.data(), anew.data());
    for (size_t i = 0; i < 2 * n0; ++i) {
      rA[i] = (anew[i] - o.A[i]) / l2; // per unit of A (the state stores lambda^2 A)
      R += 0.5 * rA[i] * rA[i];
    }
    // transpose of A' = -h^2 K * (j - mean_grain j) (solveVectorPotential's
    // correlation: out(x) = sum_p src(p) K(p - x)), seeded with rA / lambda^2
    const int N = o.N0, W = 2 * N - 1;
    for (int d = 0; d < 2; ++d) {
      std::vector<double> sbar(n0, 0.0);
      for (int py = 0; py < N; ++py)
        for (int px = 0; px < N; ++px) {
          double sum = 0.0;
          for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x)
              sum += rA[d * n0 + size_t(y) * N + x] / l2 *
                     o.filter2D[size_t(W) * (py + N - 1 - y) + (px + N - 1 - x)];
          sbar[size_t(py) * N + px] = -(o.h * o.h) * sum;
        }
      double tot = 0.0;
      for (size_t i = 0; i < n0; ++i) tot += sbar[i];
      const double mean = tot / static_cast<double>(o.numGrain);
      for (size_t i = 0; i < n0; ++i)
        jbar[d * n0 + i] = o.domainRef[i] != 0 ? (sbar[i] - mean) * jfac : 0.0;
    }
  }
  // reverse, angle by angle
  std::vector<double> seedD(2 * n0), rbarD(2 * n1), rbarJ(n1), dbar(2 * n1), esbar(n1),
      arbar(n1), sumbar(2 * n0), rho(n1, 0.0);
  std::vector<cplx> locG(size_t(Ne) * n1), locGt(size_t(Ne) * n1);
  for (int k = 0; k < o.Na; ++k) {
    rotatedOrderParameter(o, k, drot);
    energyShift(o, k, es);
    riccatiAngle(o, k, drot, es, true);
    const signed char *lab = o.frames[k].label.data();
    const float phi = static_cast<float>(o.systemAngle[k]);
    const double wk = angleWeight(k, o.Na) / static_cast<double>(o.Na);
    const double pa = polarAngle(k, o.Na);
    // seeds of the frame's Green sums: transposes of the back-rotation and
    // the extension
    std::fill(seedD.begin(), seedD.end(), 0.0);
    for (int c = 0; c < Nc; ++c) {
      const double b = basisMomentum(o, c, k) * o.coupling[c] * wk;
      for (size_t i = 0; i < 2 * n0; ++i) seedD[i] += b * rD[size_t(c)][i];
    }
    std::fill(rbarD.begin(), rbarD.end(), 0.0);
    insertRotatedAdjoint(seedD.data(), o.N0, o.N0, rbarD.data(), N1, N1, phi);
    insertRotatedAdjoint(seedD.data() + n0, o.N0, o.N0, rbarD.data() + n1, N1, N1, phi);
    extendValuesAdjoint(lab, N1, rbarD.data());
    extendValuesAdjoint(lab, N1, rbarD.data() + n1);
    std::fill(rbarJ.begin(), rbarJ.end(), 0.0);
    if (doA) {
      std::vector<double> sj(n0);
      for (size_t i = 0; i < n0; ++i)
        sj[i] = wk * (std::cos(pa) * jbar[i] + std::sin(pa) * jbar[n0 + i]);
      insertRotatedAdjoint(sj.data(), o.N0, o.N0, rbarJ.data(), N1, N1, phi);
      extendValuesAdjoint(lab, N1, rbarJ.data());
    }
    // local adjoints of gamma and gamma~: with p = g gt, q2 = 1 / (1 + p)^2,
    // gD = -2i g / (1 + p) + conj(2i gt / (1 + p)) and Re of
    // gJ = -i (1 - p) / (1 + p) (dgJ/dp = 2i q2)
    for (size_t i = 0; i < n1; ++i)
      for (int e = 0; e < Ne; ++e) {
        const size_t gi = size_t(e) * 2 * n1 + i;
        const cplx g(o.gamma[gi], o.gamma[gi + n1]), gt(o.gammat[gi], o.gammat[gi + n1]);
        const cplx Db(rbarD[i], rbarD[n1 + i]), Jb(rbarJ[i], 0.0);
        const cplx q2 = cdiv(cplx(1.0, 0.0), cmul(1.0 + cmul(g, gt), 1.0 + cmul(g, gt)));
        const cplx i2(0.0, 2.0);
        const cplx du_dg = -cmul(i2, q2), du_dgt = cmul(i2, cmul(cmul(g, g), q2));
        const cplx dv_dg = -cmul(i2, cmul(cmul(gt, gt), q2)), dv_dgt = cmul(i2, q2);
        const cplx dj_dg = cmul(i2, cmul(gt, q2)), dj_dgt = cmul(i2, cmul(g, q2));
        const double r = o.residue[e];
        locG[size_t(e) * n1 + i] = r * (cmul(std::conj(du_dg), Db) + std::conj(cmul(dv_dg, Db)) +
                                        cmul(std::conj(dj_dg), Jb));
        locGt[size_t(e) * n1 + i] =
            r * (cmul(std::conj(du_dgt), Db) + std::conj(cmul(dv_dgt, Db)) +
                 cmul(std::conj(dj_dgt), Jb));
      }
    std::fill(dbar.begin(), dbar.end(), 0.0);
    std::fill(esbar.begin(), esbar.end(), 0.0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int x = o.xmin[k]; x <= o.xmax[k]; ++x)
      for (int e = 0; e < Ne; ++e) {
        reverseColumn<true>(o, k, x, e, drot, es, rho, dbar, esbar, &locG);
        reverseColumn<false>(o, k, x, e, drot, es, rho, dbar, esbar, &locGt);
      }
    // the order parameter's rotation and extension, as in freeEnergyGradient
    std::fill(sumbar.begin(), sumbar.end(), 0.0);
    const int kk = wrapToRangeInt(-k, 0, o.Na);
    const float phiD = static_cast<float>(o.systemAngle[kk]);
    insertRotatedAdjoint(dbar.data(), N1, N1, sumbar.data(), o.N0, o.N0, phiD);
    insertRotatedAdjoint(dbar.data() + n1, N1, N1, sumbar.data() + n0, o.N0, o.N0, phiD);
    extendValuesAdjoint(o.domainRef.data(), o.N0, sumbar.data());
    extendValuesAdjoint(o.domainRef.data(), o.N0, sumbar.data() + n0);
    for (int c = 0; c < Nc; ++c) {
      const double b = basisMomentum(o, c, k);
      double *gc = grad.data() + size_t(c) * 2 * n0;
      for (size_t i = 0; i < 2 * n0; ++i) gc[i] += b * sumbar[i];
    }
    if (doA) {
      const double vf[2] = {std::cos(pa), std::sin(pa)};
      const int q = sgn(o.cfg.charge_sign);
      for (int f = 0; f < 2; ++f) {
        for (size_t i = 0; i < n1; ++i) arbar[i] = static_cast<double>(q) * vf[f] * esbar[i];
        insertRotatedAdjoint(arbar.data(), N1, N1, gradA.data() + size_t(f) * n0, o.N0, o.N0,
                             static_cast<float>(-o.systemAngle[k]));
      }
    }
  }
  // the identity part of (J - I)^T r
  for (int c = 0; c < Nc; ++c)
    for (size_t i = 0; i < 2 * n0; ++i) grad[size_t(c) * 2 * n0 + i] -= rD[size_t(c)][i];
  for (size_t i = 0; i < 2 * n0; ++i) gradA[i] -= rA[i];
  o.bOutG = keepG;
  o.bOutGt = keepGt;
  return R;
}

// ---------------------------------------------------------------------------
// Accelerators -- include/conga/accelerators/*.h
// ---------------------------------------------------------------------------
static void concatenate(const Oracle &o, bool withA,
                        const std::vector<std::vector<double>> &d,
                        const std::vector<double> &a, std::vector<double> &X) {
  const size_t n0 = o.n0sq();
  const int nf = 2 * o.Nc + (withA ? 2 : 0);
  X.assign(nf * n0, 0.0);
  for (int c = 0; c < o.Nc; ++c) {
    std::copy(d[c].begin(), d[c].begin() + n0, X.begin() + 2 * c * n0);
    std::copy(d[c].begin() + n0, d[c].end(), X.begin() + (2 * c + 1) * n0);
  }
  if (withA) std::copy(a.begin(), a.end(), X.begin() + 2 * o.Nc * n0);
}
static void split(const Oracle &o, bool withA, const std::vector<double> &X,
                  std::vector<std::vector<double>> &d, std::vector<double> &a) {
  const size_t n0 = o.n0sq();
  for (int c = 0; c < o.Nc; ++c) {
    std::copy(X.begin() + 2 * c * n0, X.begin() + (2 * c + 1) * n0,
              d[c].begin());
    std::copy(X.begin() + (2 * c + 1) * n0, X.begin() + (2 * c + 2) * n0,
              d[c].begin() + n0);
  }
  if (withA) std::copy(X.begin() + 2 * o.Nc * n0, X.end(), a.begin());
}
static double sumSq(const std::vector<double> &v) {
  double s = 0.0;
  for (double x : v) s += x * x;
  return s;
}
// Anderson.h:267-355 -- bordered normal equations solved with LU and partial
// pivoting.
static void andersonStep(Oracle &o, const std::vector<double> &G,
                         std::vector<double> &X) {
  const size_t n = X.size();
  std::vector<double> D(n);
  for (size_t i = 0; i < n; ++i) D[i] = G[i] - X[i];
  const size_t rank = o.cfg.rank > 0 ? size_t(o.cfg.rank) : 4;
  const double step = o.cfg.step_size > 0 ? o.cfg.step_size : 1.0;
  const double sq = sumSq(D);
  if (o.accD.size() == rank) {
    const size_t idx = size_t(
        std::min_element(o.accIdx.begin(), o.accIdx.end()) - o.accIdx.begin());
    o.accX[idx] = X;
    o.accD[idx] = D;
    o.accSq[idx] = sq;
    o.accIdx[idx] = size_t(o.accIter);
  } else {
    o.accX.push_back(X);
    o.accD.push_back(D);
    o.accSq.push_back(sq);
    o.accIdx.push_back(size_t(o.accIter));
  }
  const size_t m = o.accD.size(), M = m + 1;
  std::vector<double> A(M * M, 1.0), b(M, 0.0);
  double sr = 0.0;
  for (double v : o.accSq) sr += v;
  const double scale = 1.0 / (sr / static_cast<double>(m));
  const size_t n0 = o.n0sq();
  const size_t nf = n / n0;
  for (size_t i = 0; i < m; ++i) {
    A[i * M + i] = o.accSq[i] * scale;
    for (size_t j = i + 1; j < m; ++j) {
      double s = 0.0; // sumGrid: per-field sums accumulated in double
      for (size_t f = 0; f < nf; ++f) {
        double fs = 0.0;
        for (size_t q = 0; q < n0; ++q)
          fs += o.accD[i][f * n0 + q] * o.accD[j][f * n0 + q];
        s += fs;
      }
      A[i * M + j] = s * scale;
      A[j * M + i] = s * scale;
    }
  }
  A[(M - 1) * M + (M - 1)] = 0.0;
  b[M - 1] = 1.0;
  // LU with partial pivoting
  std::vector<int> piv(M);
  for (size_t i = 0; i < M; ++i) piv[i] = int(i);
  for (size_t c = 0; c < M; ++c) {
    size_t p = c;
    for (size_t r = c + 1; r < M; ++r)
      if (std::fabs(A[r * M + c]) > std::fabs(A[p * M + c])) p = r;
    if (p != c) {
      for (size_t q = 0; q < M; ++q) std::swap(A[c * M + q], A[p * M + q]);
      std::swap(b[c], b[p]);
    }
    const double dv = A[c * M + c];
    if (dv == 0.0) continue;
    for (size_t r = c + 1; r < M; ++r) {
      const double f = A[r * M + c] / dv;
      for (size_t q = c; q < M; ++q) A[r * M + q] -= f * A[c * M + q];
      b[r] -= f * b[c];
    }
  }
  std::vector<double> w(M, 0.0);
  for (size_t r = M; r-- > 0;) {
    double s = b[r];
    for (size_t q = r + 1; q < M; ++q) s -= A[r * M + q] * w[q];
    w[r] = A[r * M + r] != 0.0 ? s / A[r * M + r] : 0.0;
  }
  std::fill(X.begin(), X.end(), 0.0);
  for (size_t i = 0; i < m; ++i)
    for (size_t q = 0; q < n; ++q)
      X[q] += (o.accX[i][q] + o.accD[i][q] * step) * w[i];
  o.accIter++;
}
// CongAcc.h:271-419
static double congaccInertia(double dn, double mn, double dot, double minCos) {
  double eA = dot / (dn * mn) - minCos, xA = 1.0;
  double eB = 1.0 - minCos, xB = 0.0;
  double x = 0.5 * (xA + xB);
  bool conv = false;
  while (!conv) {
    const double y = 1.0 - x;
    const double numer = x * dot + y * dn * dn;
    const double denom =
        dn * std::sqrt(x * x * mn * mn + 2.0 * x * y * dot + y * y * dn * dn);
    const double sim = numer / denom;
    const double err = sim - minCos;
    conv = std::fabs(err) < 0.01;
    if (sgn(err) == sgn(eA)) {
      eA = err;
      xA = x;
    } else {
      eB = err;
      xB = x;
    }
    if (sgn(eA) == sgn(eB)) break;
    x = 0.5 * (xA + xB);
  }
  return x;
}
static void congaccStep(Oracle &o, const std::vector<double> &G,
                        std::vector<double> &X) {
  const size_t n = X.size(), n0 = o.n0sq(), nf = n / n0;
  const double step0 = o.cfg.step_size > 0 ? o.cfg.step_size : 0.5;
  const double fac = o.cfg.step_size_factor > 0 ? o.cfg.step_size_factor : 1.234;
  const double minCos =
      o.cfg.cos_similarity_min > 0 ? o.cfg.cos_similarity_min : 0.7;
  const double smax = o.cfg.step_size_max > 0 ? o.cfg.step_size_max : 100.0;
  const double smin = o.cfg.step_size_min > 0 ? o.cfg.step_size_min : 0.001;
  if (!o.accInit) {
    o.accMean.assign(n, 0.0);
    o.accStep.assign(nf, step0);
    o.accInertia.assign(nf, 0.0);
    o.accInit = true;
  }
  std::vector<double> D(n);
  for (size_t i = 0; i < n; ++i) D[i] = G[i] - X[i];
  for (size_t f = 0; f < nf; ++f) {
    double d2 = 0.0, m2 = 0.0, dm = 0.0;
    for (size_t q = 0; q < n0; ++q) {
      const double dv = D[f * n0 + q], mv = o.accMean[f * n0 + q];
      d2 += dv * dv;
      m2 += mv * mv;
      dm += dv * mv;
    }
    const double d = std::sqrt(d2), m = std::sqrt(m2), denom = d * m;
    if (denom > 0.0) {
      const double cs = dm / denom;
      o.accStep[f] *= std::pow(fac, cs);
      if (cs < minCos) o.accInertia[f] = congaccInertia(d, m, dm, minCos);
    }
    o.accStep[f] = std::clamp(o.accStep[f], smin, smax);
  }
  for (size_t f = 0; f < nf; ++f) {
    const double in = o.accInertia[f];
    for (size_t q = 0; q < n0; ++q) {
      const size_t i = f * n0 + q;
      const double mean = o.accMean[i] * in + D[i] * (1.0 - in);
      o.accMean[i] = mean;
      X[i] += o.accStep[f] * mean;
    }
  }
}
// BarzilaiBorwein.h:209-243
static void bbStep(Oracle &o, const std::vector<double> &G,
                   std::vector<double> &X) {
  const size_t n = X.size(), n0 = o.n0sq(), nf = n / n0;
  const double smax = o.cfg.step_size_max > 0 ? o.cfg.step_size_max : 100.0;
  const double smin = o.cfg.step_size_min > 0 ? o.cfg.step_size_min : 0.001;
  if (!o.accInit) {
    o.bbStep = o.cfg.step_size > 0 ? o.cfg.step_size : 1.0;
    o.accInit = true;
  }
  std::vector<double> D(n);
  for (size_t i = 0; i < n; ++i) D[i] = G[i] - X[i];
  if (o.accIter > 0) {
    std::vector<double> dD(n);
    for (size_t i = 0; i < n; ++i) dD[i] = D[i] - o.accPrev[i];
    double dot = 0.0;
    for (size_t f = 0; f < nf; ++f) {
      double fs = 0.0;
      for (size_t q = 0; q < n0; ++q) fs += o.accPrev[f * n0 + q] * dD[f * n0 + q];
      dot += fs;
    }
    dot = std::fabs(dot);
    const double sc = (o.accIter % 2 == 0) ? dot / sumSq(dD) : sumSq(o.accPrev) / dot;
    o.bbStep *= sc;
  }
  o.bbStep = std::clamp(o.bbStep, smin, smax);
  o.accPrev = D;
  for (size_t i = 0; i < n; ++i) X[i] += D[i] * o.bbStep;
  o.accIter++;
}
// Polyak.h:216-265
static void polyakStep(Oracle &o, const std::vector<double> &G,
                       std::vector<double> &X) {
  const size_t n = X.size();
  const double step = o.cfg.step_size > 0 ? o.cfg.step_size : 0.5;
  const double drag = o.cfg.drag > 0 ? o.cfg.drag : 0.5;
  if (!o.accInit) {
    o.accVel.assign(n, 0.0);
    o.accInit = true;
  }
  for (size_t i = 0; i < n; ++i) {
    const double d = G[i] - X[i];
    const double v = o.accVel[i] * (1.0 - drag) + d * step;
    o.accVel[i] = v;
    X[i] += v;
  }
}
// Picard.h:133-138 (applied per component and to A separately)
static void picardStep(Oracle &o, const std::vector<double> &G,
                       std::vector<double> &X) {
  const double step = o.cfg.step_size > 0 ? o.cfg.step_size : 1.0;
  for (size_t i = 0; i < X.size(); ++i) X[i] += (G[i] - X[i]) * step;
}

// Broyden's second method, restarted limited memory.  No counterpart in the
// reference's accelerators/: it restates scipy.optimize.broyden2 (scipy
// 1.15, optimize/_nonlin.py: BroydenSecond._update, LowRankMatrix.matvec and
// restart_reduce, nonlin_solve without line search) on F(x) = g(x) - x with
// alpha = step size, the dot products summed field by field like Anderson's.
static double fieldDot(const Oracle &o, const std::vector<double> &a,
                       const std::vector<double> &b) {
  const size_t n0 = o.n0sq(), nf = a.size() / n0;
  double s = 0.0;
  for (size_t f = 0; f < nf; ++f) {
    double fs = 0.0;
    for (size_t q = 0; q < n0; ++q) fs += a[f * n0 + q] * b[f * n0 + q];
    s += fs;
  }
  return s;
}
static void broydenStep(Oracle &o, const std::vector<double> &G, std::vector<double> &X) {
  const size_t n = X.size();
  const size_t rank = o.cfg.rank > 0 ? size_t(o.cfg.rank) : 8;
  const double alpha = o.cfg.step_size > 0 ? o.cfg.step_size : 1.0;
  std::vector<double> F(n);
  for (size_t i = 0; i < n; ++i) F[i] = G[i] - X[i];
  // H v = -alpha v + sum_i c_i (d_i . v), terms added in the order stored
  auto applyH = [&](const std::vector<double> &v) {
    std::vector<double> w(n);
    for (size_t i = 0; i < n; ++i) w[i] = -alpha * v[i];
    for (size_t k = 0; k < o.brC.size(); ++k) {
      const double a = fieldDot(o, o.brD[k], v);
      for (size_t i = 0; i < n; ++i) w[i] += a * o.brC[k][i];
    }
    return w;
  };
  if (o.accIter > 0) {
    std::vector<double> dF(n), dx(n);
    for (size_t i = 0; i < n; ++i) {
      dF[i] = F[i] - o.brLastF[i];
      dx[i] = X[i] - o.brLastX[i];
    }
    // restart_reduce(rank - 1): drop every pair once more than rank - 1 are held
    if (o.brC.size() > rank - 1) {
      o.brC.clear();
      o.brD.clear();
    }
    const double dfn2 = fieldDot(o, dF, dF);
    if (dfn2 > 0.0) {
      const std::vector<double> hdf = applyH(dF);
      std::vector<double> c(n), d(n);
      for (size_t i = 0; i < n; ++i) {
        c[i] = dx[i] - hdf[i];
        d[i] = dF[i] / dfn2;
      }
      o.brC.push_back(std::move(c));
      o.brD.push_back(std::move(d));
    }
  }
  o.brLastX = X;
  o.brLastF = F;
  const std::vector<double> hf = applyH(F);
  for (size_t i = 0; i < n; ++i) X[i] -= hf[i];
  o.accIter++;
}

static void acceleratorStep(Oracle &o, bool doOP, bool doA) {
  std::vector<double> X, G;
  const std::vector<double> empty;
  if (doOP) {
    concatenate(o, doA, o.delta, o.A, X);
    concatenate(o, doA, o.deltaNew, o.Anew, G);
  } else {
    X = o.A;
    G = o.Anew;
  }
  switch (o.cfg.accelerator) {
  case OC_PICARD: picardStep(o, G, X); break;
  case OC_ANDERSON: andersonStep(o, G, X); break;
  case OC_CONGACC: congaccStep(o, G, X); break;
  case OC_BB: bbStep(o, G, X); break;
  case OC_POLYAK: polyakStep(o, G, X); break;
  case OC_BROYDEN: broydenStep(o, G, X); break;
  }
  if (doOP)
    split(o, doA, X, o.delta, o.A);
  else
    o.A = X;
}

// Context.h:391-447
static void contextCompute(Oracle &o) {
  const bool doA = o.cfg.compute_current && o.lambda > 0.0;
  if (!o.firstCompute) acceleratorStep(o, true, doA);
  integrate(o, false);
  if (doA) {
    if (o.filter2D.empty())
      makeFilter2D(o.N0, o.h, o.filter2D, std::max(o.cfg.vector_potential_error, 0.0));
    solveVectorPotential(o.N0, o.h, o.numGrain, true, o.filter2D,
                         o.current.data(), o.Anew.data());
    o.resA = residualArrays(o.A.data(), o.Anew.data(), o.A.size(), o.cfg.norm);
  }
  o.numIterations++;
  o.firstCompute = false;
}
// Context.h:462-497
static double contextResidual(const Oracle &o) {
  double r = 0.0;
  if (o.cfg.compute_current && o.lambda > 0.0) r = o.resA;
  r = std::max(o.resBoundary, r);
  r = std::max(o.resOP, r);
  if (o.cfg.compute_current) r = std::max(o.resJ, r);
  if (o.cfg.compute_free_energy) r = std::max(o.resFE, r);
  return r;
}
static bool contextConverged(const Oracle &o) {
  bool ok = contextResidual(o) < o.cfg.convergence_criterion;
  ok = ok && o.convOP;
  if (o.cfg.compute_current) ok = ok && o.convJ;
  if (o.cfg.compute_free_energy) ok = ok && o.convFE;
  return ok;
}

// ---------------------------------------------------------------------------
// Keldysh spectroscopy (src/postprocess.cu:64-472): the LDOS at real
// energies E + i delta, energy block by energy block.  The context holds no
// order-parameter, current or vector-potential compute object, so there is
// no accelerator step and no Poisson solve (Context.h:391-447).
// ---------------------------------------------------------------------------
// IntegrationIteratorKeldysh::initEnergies (:115-150)
static double keldyshEnergy(const Oracle &o, int i) {
  const int n = o.kel.numEnergies;
  const double step = n > 1 ? (o.kel.eMax - o.kel.eMin) / static_cast<double>(n - 1) : 0.0;
  return o.kel.eMin + step * static_cast<double>(i);
}
// A new energy interval: the boundary storage for its energies and the
// homogeneous initial condition (IntegrationIteratorKeldysh.h:184-232,274-293).
static void keldyshBeginBlock(Oracle &o) {
  const int start = o.kel.blockStart;
  o.Ne = std::min(o.kel.perBlock, o.kel.numEnergies - start);
  o.energyRe.resize(size_t(o.Ne));
  o.energyIm.assign(size_t(o.Ne), o.kel.broadening);
  o.residue.assign(size_t(o.Ne), 1.0);
  for (int e = 0; e < o.Ne; ++e) o.energyRe[size_t(e)] = keldyshEnergy(o, start + e);
  o.bInG.assign(size_t(2 * o.nInG * o.Ne), 0.0);
  o.bOutG.assign(size_t(2 * o.nOutG * o.Ne), 0.0);
  o.bInGt.assign(size_t(2 * o.nInGt * o.Ne), 0.0);
  o.bOutGt.assign(size_t(2 * o.nOutGt * o.Ne), 0.0);
  o.gamma.assign(size_t(o.Ne) * 2 * o.n1sq(), 0.0);
  o.gammat.assign(size_t(o.Ne) * 2 * o.n1sq(), 0.0);
  o.kel.interval.assign(size_t(o.Ne) * o.n0sq(), 0.0);
  setBoundaryInitialCondition(o);
  o.kel.partial = false;
  o.kel.lastBlock = start;
}
static bool keldyshDone(const Oracle &o) { return o.kel.blockStart >= o.kel.numEnergies; }

// IntegrationIteratorKeldysh::compute (:235-388) with ComputeLDOS
// (ComputeLDOS.h:110-192), then Context::compute / burnIn bookkeeping.
static void keldyshCompute(Oracle &o, bool burnin) {
  if (!keldyshDone(o)) {
    if (o.kel.lastBlock != o.kel.blockStart) keldyshBeginBlock(o);
    const size_t n0 = o.n0sq(), n1 = o.n1sq();
    std::vector<double> drot, es;
    if (burnin) {
      for (int k = 0; k < o.Na; ++k) {
        rotatedOrderParameter(o, k, drot);
        energyShift(o, k, es);
        riccatiAngle(o, k, drot, es, false);
      }
      boundaryCondition(o);
    } else {
      std::fill(o.kel.interval.begin(), o.kel.interval.end(), 0.0);
      std::vector<double> gim(n1);
      for (int k = 0; k < o.Na; ++k) {
        rotatedOrderParameter(o, k, drot);
        energyShift(o, k, es);
        riccatiAngle(o, k, drot, es, true);
        // ComputeLDOS::computeGreensAngle / integrateGreensAngle (:126-151):
        // Greens per energy, extended at the boundary, times the integrand
        // factor, imaginary part rotated back additively.
        const signed char *lab = o.frames[size_t(k)].label.data();
        const double wk = angleWeight(k, o.Na) / static_cast<double>(o.Na);
        const float phi = static_cast<float>(o.systemAngle[size_t(k)]);
        for (int e = 0; e < o.Ne; ++e) {
#pragma omp parallel for schedule(static)
          for (long long i = 0; i < (long long)n1; ++i) {
            const size_t gi = size_t(e) * 2 * n1 + size_t(i);
            const cplx ga(o.gamma[gi], o.gamma[gi + n1]);
            const cplx gt(o.gammat[gi], o.gammat[gi + n1]);
            gim[size_t(i)] = greensG(ga, gt).imag();
          }
          extendValues(lab, o.N1, gim.data());
          for (double &v : gim) v *= wk;
          insertRotated(o.kel.interval.data() + size_t(e) * n0, o.N0, o.N0, gim.data(), o.N1,
                        o.N1, phi, true);
        }
      }
      boundaryCondition(o);
      // ComputeLDOS::computePropertyFromGreens (:153-192)
      for (double &v : o.kel.interval) v *= -1.0;
      for (int e = 0; e < o.Ne; ++e) zeroOutside(o, o.kel.interval.data() + size_t(e) * n0);
      double *slot = o.kel.all.data() + size_t(o.kel.blockStart) * n0;
      const size_t len = size_t(o.Ne) * n0;
      o.kel.residual = residualArrays(slot, o.kel.interval.data(), len, o.cfg.norm);
      std::copy(o.kel.interval.begin(), o.kel.interval.begin() + std::ptrdiff_t(len), slot);
      o.kel.partial = o.kel.residual < o.cfg.convergence_criterion;
      const bool boundaryConverged = o.resBoundary < o.cfg.convergence_criterion;
      if (o.kel.partial && boundaryConverged) {
        o.kel.blockStart += o.kel.perBlock;
        if (keldyshDone(o)) o.kel.converged = true;
      }
    }
  }
  o.numIterations++;
}
// Context::getResidual / isPartiallyConverged / isConverged (Context.h:462-511)
static double keldyshResidual(const Oracle &o) { return std::max(o.resBoundary, o.kel.residual); }
static bool keldyshPartiallyConverged(const Oracle &o) {
  return keldyshResidual(o) < o.cfg.convergence_criterion && o.kel.partial;
}
static bool keldyshConverged(const Oracle &o) {
  return keldyshResidual(o) < o.cfg.convergence_criterion && o.kel.converged;
}
// run_iteration.h:85-141
static bool keldyshIterations(Oracle &o, bool compute, int mn, int mx) {
  if (mn == 0 && mx == 0) return false;
  bool conv = false;
  for (int i = 1;; ++i) {
    if (compute) {
      keldyshCompute(o, false);
      conv = keldyshPartiallyConverged(o);
    } else {
      keldyshCompute(o, true);
      conv = o.resBoundary < o.cfg.convergence_criterion;
    }
    const bool minReached = i >= mn;
    const bool maxReached = mx < 0 ? false : i >= mx;
    if ((conv && minReached) || maxReached) break;
  }
  return conv;
}
// runPostprocess (run_simulation.h:44-108)
static bool keldyshRun(Oracle &o, int nb, int mn, int mx) {
  bool converged = false;
  while (!converged) {
    keldyshIterations(o, false, nb, nb);
    const bool blockConverged = keldyshIterations(o, true, mn, mx);
    converged = keldyshConverged(o);
    if (!blockConverged) break;
  }
  return converged;
}

// ---------------------------------------------------------------------------
// Construction: simulation.cu:64-463 + Context::initialize (Context.h:285-375)
// ---------------------------------------------------------------------------
static Oracle *build(const oc_config &cfg) {
  auto o = std::make_unique<Oracle>();
  o->cfg = cfg;
  o->T = cfg.temperature;
  o->lambda = cfg.penetration_depth;
  o->Na = cfg.num_fermi_momenta;
  o->Nc = cfg.num_components;
  if (o->Nc < 1 || o->Nc > OC_MAX_COMPONENTS) {
    g_error = "invalid number of order-parameter components";
    return nullptr;
  }
  // geometry scaling (geometry_utils.h:41-175) and components (:193-325)
  double width;
  std::vector<Component> comps;
  if (cfg.geometry_internal) {
    width = cfg.grain_width;
  } else {
    double xMax = std::numeric_limits<double>::min();
    double xMin = std::numeric_limits<double>::max();
    double yMax = std::numeric_limits<double>::min();
    double yMin = std::numeric_limits<double>::max();
    for (int i = 0; i < cfg.num_geometry; ++i) {
      if (!cfg.geometry_add[i]) continue;
      const double *p = cfg.geometry_param[i];
      if (cfg.geometry_type[i] == OC_DISC) {
        xMax = std::max(xMax, p[0] + p[2]);
        xMin = std::min(xMin, p[0] - p[2]);
        yMax = std::max(yMax, p[1] + p[2]);
        yMin = std::min(yMin, p[1] - p[2]);
      } else if (cfg.geometry_type[i] == OC_POLYGON) {
        for (int v = 0; v < cfg.polygon_num_vertices[i]; ++v) {
          const double x = cfg.polygon_vertices_x[cfg.polygon_vertex_offset[i] + v];
          const double y = cfg.polygon_vertices_y[cfg.polygon_vertex_offset[i] + v];
          xMax = std::max(xMax, x);
          xMin = std::min(xMin, x);
          yMax = std::max(yMax, y);
          yMin = std::min(yMin, y);
        }
      } else {
        const double side = p[4], cx = p[0], cy = p[1];
        const double rot = 2.0 * M_PI * p[3];
        const double ne = static_cast<double>(static_cast<int>(p[2]));
        const double step = 2.0 * M_PI / ne;
        const double radius =
            side / std::hypot(std::cos(step) - 1.0, std::sin(step));
        const double off = 0.5 * step + rot;
        for (int v = 0; v < ne; ++v) {
          const double a = off + step * static_cast<double>(v);
          const double x = cx + radius * std::cos(a);
          const double y = cy + radius * std::sin(a);
          xMax = std::max(xMax, x);
          xMin = std::min(xMin, x);
          yMax = std::max(yMax, y);
          yMin = std::min(yMin, y);
        }
      }
    }
    const double xe = xMax - xMin, ye = yMax - yMin;
    const double offX = xMin + 0.5 * xe, offY = yMin + 0.5 * ye;
    width = std::max(xe, ye);
    // components in internal units (before grain-fraction scaling)
    for (int i = 0; i < cfg.num_geometry; ++i) {
      Component c;
      c.add = cfg.geometry_add[i] != 0;
      const double *p = cfg.geometry_param[i];
      if (cfg.geometry_type[i] == OC_DISC) {
        c.type = OC_DISC;
        c.cx = (p[0] - offX) / width;
        c.cy = (p[1] - offY) / width;
        c.radius = p[2] / width;
      } else if (cfg.geometry_type[i] == OC_POLYGON) {
        c.type = OC_POLYGON;
        for (int v = 0; v < cfg.polygon_num_vertices[i]; ++v) {
          c.vx.push_back(
              (cfg.polygon_vertices_x[cfg.polygon_vertex_offset[i] + v] - offX) /
              width);
          c.vy.push_back(
              (cfg.polygon_vertices_y[cfg.polygon_vertex_offset[i] + v] - offY) /
              width);
        }
      } else {
        c.type = OC_REGULAR_POLYGON;
        c.cx = (p[0] - offX) / width;
        c.cy = (p[1] - offY) / width;
        c.radius = p[4] / width; // side length for now
        c.vx.assign(1, 2.0 * M_PI * p[3]);
        c.vy.assign(1, p[2]);
      }
      comps.push_back(c);
    }
  }
  lattice(width, cfg.points_per_coherence_length, o->N0, o->N1,
          o->grainFraction);
  o->h = 1.0 / cfg.points_per_coherence_length;
  const double gf = o->grainFraction;
  if (cfg.geometry_internal) {
    for (int i = 0; i < cfg.num_geometry; ++i) {
      Component c;
      c.add = cfg.geometry_add[i] != 0;
      const double *p = cfg.geometry_param[i];
      if (cfg.geometry_type[i] == OC_DISC) {
        // DiscGeometry.h:249-261
        c.type = OC_DISC;
        c.radius = p[2] * gf;
        c.cx = p[0] * gf;
        c.cy = p[1] * gf;
      } else if (cfg.geometry_type[i] == OC_POLYGON) {
        // PolygonGeometry.h:61-77: vertices scaled by the grain fraction
        c.type = OC_POLYGON;
        for (int v = 0; v < cfg.polygon_num_vertices[i]; ++v) {
          c.vx.push_back(cfg.polygon_vertices_x[cfg.polygon_vertex_offset[i] + v] * gf);
          c.vy.push_back(cfg.polygon_vertices_y[cfg.polygon_vertex_offset[i] + v] * gf);
        }
      } else {
        // PolygonGeometry.h:265-291
        c.type = OC_POLYGON;
        const int nv = static_cast<int>(p[2]);
        const double step = 2.0 * M_PI / static_cast<double>(nv);
        const double rfs = p[4] / std::hypot(std::cos(step) - 1.0, std::sin(step));
        const double radius = p[4] < 0.0 ? MAXIMUM_COORDINATE : rfs;
        regularPolygonVertices(nv, radius * gf, p[3], p[0] * gf, p[1] * gf, c.vx,
                               c.vy);
      }
      o->comps.push_back(c);
    }
  } else {
    for (Component c : comps) {
      if (c.type == OC_DISC) {
        c.radius *= gf;
        c.cx *= gf;
        c.cy *= gf;
      } else if (c.type == OC_POLYGON) {
        for (double &v : c.vx) v *= gf;
        for (double &v : c.vy) v *= gf;
      } else {
        const int nv = static_cast<int>(c.vy[0]);
        const double rot = c.vx[0];
        const double side = c.radius;
        const double step = 2.0 * M_PI / static_cast<double>(nv);
        const double rfs = side / std::hypot(std::cos(step) - 1.0, std::sin(step));
        const double radius = side < 0.0 ? MAXIMUM_COORDINATE : rfs;
        c.type = OC_POLYGON;
        regularPolygonVertices(nv, radius * gf, rot, c.cx * gf, c.cy * gf, c.vx,
                               c.vy);
      }
      o->comps.push_back(c);
    }
  }
  // energies
  o->Ne = numOzakiEnergies(o->T, cfg.energy_cutoff);
  std::vector<double> poles, res;
  ozakiPoles(o->Ne, poles, res);
  o->energyIm.resize(o->Ne);
  o->energyRe.assign(o->Ne, 0.0);
  o->residue = res;
  for (int i = 0; i < o->Ne; ++i)
    o->energyIm[i] = poles[i] * o->T / static_cast<double>(2.0 * M_PI);
  // Fermi surface
  o->systemAngle.resize(o->Na);
  o->polar.resize(o->Na);
  for (int k = 0; k < o->Na; ++k) {
    o->polar[k] = polarAngle(k, o->Na);
    o->systemAngle[k] = o->polar[k] - static_cast<double>(0.5 * M_PI);
  }
  // geometry frames: GeometryGroup::create(-systemAngle(k))
  o->frames.resize(o->Na);
#pragma omp parallel for schedule(dynamic, 1)
  for (int k = 0; k < o->Na; ++k)
    createGroup(o->comps, -o->systemAngle[k], o->N0, o->N1, o->frames[k]);
  o->domainRef = createDomain(o->comps, o->N0);
  o->numGrain = 0;
  for (signed char v : o->domainRef) o->numGrain += v > 0 ? 1 : 0;
  // order parameter
  for (int c = 0; c < o->Nc; ++c) {
    o->symmetry.push_back(cfg.symmetry[c]);
    o->tc.push_back(cfg.critical_temperature[c]);
    // BasisFunction::normalize (OrderParameterBasis.h:223-250)
    double avg = 0.0;
    for (int k = 0; k < o->Na; ++k) {
      const double phi = polarAngle(k, o->Na);
      const double b = basisRaw(cfg.symmetry[c], std::cos(phi), std::sin(phi));
      avg += b * b * (angleWeight(k, o->Na) / static_cast<double>(o->Na));
    }
    o->basisNorm.push_back(1.0 / std::sqrt(avg));
    // OrderParameterComponent::initCouplingConstant (:156-181)
    double es = 0.0;
    for (int i = 0; i < o->Ne; ++i) es += o->residue[i] / o->energyIm[i];
    es *= o->T;
    o->coupling.push_back(0.5 * o->T / (std::log(o->T / o->tc[c]) + es));
    o->delta.emplace_back(2 * o->n0sq(), 0.0);
    o->deltaNew.emplace_back(2 * o->n0sq(), 0.0);
  }
  o->A.assign(2 * o->n0sq(), 0.0);
  o->Anew.assign(2 * o->n0sq(), 0.0);
  o->current.assign(2 * o->n0sq(), 0.0);
  o->currentAcc.assign(4 * o->n0sq(), 0.0);
  o->feGrid.assign(2 * o->n0sq(), 0.0);
  initBoundaryStorage(*o);
  o->gamma.assign(size_t(o->Ne) * 2 * o->n1sq(), 0.0);
  o->gammat.assign(size_t(o->Ne) * 2 * o->n1sq(), 0.0);
  // the Poisson filter is built on the first solve that needs it
  return o.release();
}

} // namespace oracle

using namespace oracle;

extern "C" {

void *oc_create(const oc_config *config) {
  try {
#ifdef _OPENMP
    if (config->num_threads > 0) omp_set_num_threads(config->num_threads);
#endif
    return build(*config);
  } catch (const std::exception &e) {
    g_error = e.what();
    return nullptr;
  }
}
void oc_destroy(void *h) { delete static_cast<Oracle *>(h); }
const char *oc_last_error(void) { return g_error.c_str(); }

long long oc_get_int(void *hh, const char *key) {
  const Oracle &o = *static_cast<Oracle *>(hh);
  const std::string k(key);
  if (k == "N0") return o.N0;
  if (k == "N1") return o.N1;
  if (k == "num_energies") return o.Ne;
  if (k == "num_angles") return o.Na;
  if (k == "num_components") return o.Nc;
  if (k == "num_grain_sites") return o.numGrain;
  if (k == "num_storage_in_gamma") return o.nInG;
  if (k == "num_storage_out_gamma") return o.nOutG;
  if (k == "num_storage_in_gammatilde") return o.nInGt;
  if (k == "num_storage_out_gammatilde") return o.nOutGt;
  if (k == "num_iterations") return o.numIterations;
  if (k == "converged") return contextConverged(o) ? 1 : 0;
  if (k == "ldos_block_start") return o.kel.blockStart;
  if (k == "ldos_converged") return keldyshConverged(o) ? 1 : 0;
  if (k == "ldos_partially_converged") return keldyshPartiallyConverged(o) ? 1 : 0;
  if (k == "traj_steps") {
    long long s = 0;
    for (int a = 0; a < o.Na; ++a)
      for (signed char v : o.frames[a].label) s += v > 0 ? 1 : 0;
    return 2 * s * o.Ne;
  }
  if (k.rfind("xmin_", 0) == 0) return o.xmin[std::stoi(k.substr(5))];
  if (k.rfind("xmax_", 0) == 0) return o.xmax[std::stoi(k.substr(5))];
  if (k.rfind("ymin_", 0) == 0) return o.ymin[std::stoi(k.substr(5))];
  if (k.rfind("ymax_", 0) == 0) return o.ymax[std::stoi(k.substr(5))];
  g_error = "unknown key " + k;
  return -1;
}

double oc_get_double(void *hh, const char *key) {
  const Oracle &o = *static_cast<Oracle *>(hh);
  const std::string k(key);
  if (k == "residual") return o.kel.on ? keldyshResidual(o) : contextResidual(o);
  if (k == "residual_ldos") return o.kel.residual;
  if (k == "residual_boundary") return o.resBoundary;
  if (k == "residual_order_parameter") return o.resOP;
  if (k == "residual_current") return o.resJ;
  if (k == "residual_free_energy") return o.resFE;
  if (k == "residual_vector_potential") return o.resA;
  if (k == "free_energy") return o.feTotal;
  if (k == "free_energy_per_area") return o.fePerArea;
  if (k == "magnetic_free_energy") return o.feMag;
  if (k == "magnetic_free_energy_per_area") return o.feMagPerArea;
  if (k == "grain_fraction") return o.grainFraction;
  if (k == "h") return o.h;
  if (k.rfind("coupling_", 0) == 0) return o.coupling[std::stoi(k.substr(9))];
  if (k.rfind("basis_norm_", 0) == 0) return o.basisNorm[std::stoi(k.substr(11))];
  if (k.rfind("delta_bulk_", 0) == 0) {
    const int c = std::stoi(k.substr(11));
    const double s = bulkGap(o.T);
    const int sym = o.symmetry[c];
    return sym == OC_SWAVE ? s : s * (D_WAVE_BULK / S_WAVE_BULK);
  }
  if (k.rfind("energy_", 0) == 0) return o.energyIm[std::stoi(k.substr(7))];
  if (k.rfind("residue_", 0) == 0) return o.residue[std::stoi(k.substr(8))];
  if (k.rfind("system_angle_", 0) == 0)
    return o.systemAngle[std::stoi(k.substr(13))];
  g_error = "unknown key " + k;
  return std::nan("");
}

long long oc_get_array(void *hh, const char *name, double *out,
                       long long cap) {
  const Oracle &o = *static_cast<Oracle *>(hh);
  const std::string k(name);
  std::vector<double> tmp;
  const std::vector<double> *src = nullptr;
  if (k == "delta" || k == "delta_new") {
    for (int c = 0; c < o.Nc; ++c) {
      const auto &d = (k == "delta") ? o.delta[c] : o.deltaNew[c];
      tmp.insert(tmp.end(), d.begin(), d.end());
    }
    src = &tmp;
  } else if (k == "current") {
    src = &o.current;
  } else if (k == "vector_potential") {
    src = &o.A;
  } else if (k == "vector_potential_new") {
    src = &o.Anew;
  } else if (k == "free_energy_grid") {
    src = &o.feGrid;
  } else if (k == "ldos") {
    src = &o.kel.all;
  } else if (k == "domain") {
    tmp.assign(o.domainRef.begin(), o.domainRef.end());
    src = &tmp;
  } else if (k == "boundary_in_gamma") {
    src = &o.bInG;
  } else if (k == "boundary_out_gamma") {
    src = &o.bOutG;
  } else if (k == "boundary_in_gammatilde") {
    src = &o.bInGt;
  } else if (k == "boundary_out_gammatilde") {
    src = &o.bOutGt;
  } else {
    g_error = "unknown array " + k;
    return -1;
  }
  if ((long long)src->size() > cap) {
    g_error = "capacity too small";
    return -1;
  }
  std::copy(src->begin(), src->end(), out);
  return (long long)src->size();
}

int oc_set_order_parameter(void *hh, int c, const double *re,
                           const double *im) {
  Oracle &o = *static_cast<Oracle *>(hh);
  if (c < 0 || c >= o.Nc) return -1;
  const size_t n0 = o.n0sq();
  std::copy(re, re + n0, o.delta[c].begin());
  std::copy(im, im + n0, o.delta[c].begin() + n0);
  zeroOutside(o, o.delta[c].data());
  zeroOutside(o, o.delta[c].data() + n0);
  return 0;
}
int oc_set_vector_potential(void *hh, const double *ax, const double *ay) {
  Oracle &o = *static_cast<Oracle *>(hh);
  const size_t n0 = o.n0sq();
  const double s = o.lambda > 0.0 ? o.lambda * o.lambda : 0.0;
  for (size_t i = 0; i < n0; ++i) {
    o.A[i] = ax[i] * s;
    o.A[n0 + i] = ay[i] * s;
  }
  return 0;
}
int oc_set_array(void *hh, const char *name, const double *in,
                 long long count) {
  Oracle &o = *static_cast<Oracle *>(hh);
  const std::string k(name);
  std::vector<double> *dst = nullptr;
  if (k == "boundary_in_gamma") dst = &o.bInG;
  else if (k == "boundary_in_gammatilde") dst = &o.bInGt;
  else if (k == "vector_potential") dst = &o.A;
  if (!dst || (long long)dst->size() != count) return -1;
  std::copy(in, in + count, dst->begin());
  if (k.rfind("boundary_in", 0) == 0) o.boundaryInit = true;
  return 0;
}
int oc_initial_bulk(void *hh, double phase) {
  Oracle &o = *static_cast<Oracle *>(hh);
  const size_t n0 = o.n0sq();
  for (int c = 0; c < o.Nc; ++c) {
    const double s = bulkGap(o.T);
    const double db =
        o.symmetry[c] == OC_SWAVE ? s : s * (D_WAVE_BULK / S_WAVE_BULK);
    for (size_t i = 0; i < n0; ++i) {
      o.delta[c][i] = db * std::cos(phase);
      o.delta[c][n0 + i] = db * std::sin(phase);
    }
    zeroOutside(o, o.delta[c].data());
    zeroOutside(o, o.delta[c].data() + n0);
  }
  return 0;
}
int oc_get_geometry(void *hh, int k, signed char *labels, double *dist,
                    double *nx, double *ny) {
  const Oracle &o = *static_cast<Oracle *>(hh);
  if (k < 0 || k >= o.Na) return -1;
  const FrameGeometry &g = o.frames[k];
  std::copy(g.label.begin(), g.label.end(), labels);
  if (dist) std::copy(g.dist.begin(), g.dist.end(), dist);
  if (nx) std::copy(g.nx.begin(), g.nx.end(), nx);
  if (ny) std::copy(g.ny.begin(), g.ny.end(), ny);
  return 0;
}
int oc_compute(void *hh) {
  contextCompute(*static_cast<Oracle *>(hh));
  return 0;
}
int oc_burnin(void *hh) {
  Oracle &o = *static_cast<Oracle *>(hh);
  integrate(o, true);
  o.numIterations++;
  return 0;
}
// run_simulation.h:124-187 with run_iteration.h:85-141 (save_frequency < 0)
int oc_run(void *hh, int nb, int mn, int mx) {
  Oracle &o = *static_cast<Oracle *>(hh);
  if (nb > 0) {
    for (int i = 1;; ++i) {
      integrate(o, true);
      o.numIterations++;
      const bool conv = o.resBoundary < o.cfg.convergence_criterion;
      if ((conv && i >= nb) || i >= nb) break;
    }
  }
  if (mn == 0 && mx == 0) return 0;
  bool conv = false;
  for (int i = 1;; ++i) {
    contextCompute(o);
    conv = contextConverged(o);
    const bool minReached = i >= mn;
    const bool maxReached = mx < 0 ? false : i >= mx;
    if ((conv && minReached) || maxReached) break;
  }
  return conv ? 1 : 0;
}
// Turns the context into the postprocess context of src/postprocess.cu:
// energies min..max (num) with broadening, blocks of per_block energies
// (< 1: all).  The order parameter and vector potential are the current ones.
int oc_ldos_setup(void *hh, double emin, double emax, int num, double broadening,
                  int perBlock) {
  Oracle &o = *static_cast<Oracle *>(hh);
  if (num < 1) {
    g_error = "num_energies must be positive";
    return -1;
  }
  o.kel = Oracle::Spectroscopy{};
  o.kel.on = true;
  o.kel.eMin = emin;
  o.kel.eMax = emax;
  o.kel.numEnergies = num;
  o.kel.broadening = broadening;
  o.kel.perBlock = perBlock < 1 ? num : std::min(perBlock, num);
  o.kel.all.assign(size_t(num) * o.n0sq(), 0.0);
  o.resBoundary = std::numeric_limits<double>::max();
  o.numIterations = 0;
  return 0;
}
int oc_ldos_compute(void *hh, int burnin) {
  Oracle &o = *static_cast<Oracle *>(hh);
  if (!o.kel.on) {
    g_error = "oc_ldos_setup was not called";
    return -1;
  }
  keldyshCompute(o, burnin != 0);
  return 0;
}
int oc_ldos_run(void *hh, int nb, int mn, int mx) {
  Oracle &o = *static_cast<Oracle *>(hh);
  if (!o.kel.on) {
    g_error = "oc_ldos_setup was not called";
    return -1;
  }
  return keldyshRun(o, nb, mn, mx) ? 1 : 0;
}

int oc_free_energy_gradient(void *hh, double *grad, double *free_energy) {
  try {
    Oracle &o = *static_cast<Oracle *>(hh);
    std::vector<double> g;
    const double f = freeEnergyGradient(o, g);
    std::copy(g.begin(), g.end(), grad);
    if (free_energy) *free_energy = f;
    return 0;
  } catch (const std::exception &e) {
    g_error = e.what();
    return -1;
  }
}

int oc_plain_map(void *hh, double *deltaNew, double *aNew) {
  // G(u) of one plain iteration with the boundary in-values held fixed:
  // the gap equation's order parameter (zero outside) and the vector
  // potential in the setters' units; the out-values restored
  try {
    Oracle &o = *static_cast<Oracle *>(hh);
    if (!o.boundaryInit) setBoundaryInitialCondition(o);
    const std::vector<double> keepG = o.bOutG, keepGt = o.bOutGt;
    const size_t n0 = o.n0sq();
    const bool doA = o.cfg.compute_current && o.lambda > 0.0;
    std::vector<double> drot, es, gD, gJ, gF, tmp(2 * n0), jacc(2 * n0, 0.0), j(2 * n0, 0.0),
        anew(2 * n0, 0.0);
    std::fill(deltaNew, deltaNew + size_t(o.Nc) * 2 * n0, 0.0);
    for (int k = 0; k < o.Na; ++k) {
      rotatedOrderParameter(o, k, drot);
      energyShift(o, k, es);
      riccatiAngle(o, k, drot, es, true);
      greensAngle(o, drot, gD, gJ, gF);
      const signed char *lab = o.frames[k].label.data();
      const float phi = static_cast<float>(o.systemAngle[k]);
      const double wk = angleWeight(k, o.Na) / static_cast<double>(o.Na);
      const double pa = polarAngle(k, o.Na);
      extendValues(lab, o.N1, gD.data());
      extendValues(lab, o.N1, gD.data() + o.n1sq());
      std::fill(tmp.begin(), tmp.end(), 0.0);
      insertRotated(tmp.data(), o.N0, o.N0, gD.data(), o.N1, o.N1, phi, false);
      insertRotated(tmp.data() + n0, o.N0, o.N0, gD.data() + o.n1sq(), o.N1, o.N1, phi, false);
      for (int c = 0; c < o.Nc; ++c) {
        const double b = basisMomentum(o, c, k) * wk;
        for (size_t i = 0; i < 2 * n0; ++i) deltaNew[size_t(c) * 2 * n0 + i] += tmp[i] * b;
      }
      if (doA) {
        extendValues(lab, o.N1, gJ.data());
        for (int f = 0; f < 2; ++f) {
          std::vector<double> t2(gJ.begin(), gJ.begin() + o.n1sq());
          const double fac = (f == 0 ? std::cos(pa) : std::sin(pa)) * wk;
          for (double &v : t2) v *= fac;
          insertRotated(jacc.data() + size_t(f) * n0, o.N0, o.N0, t2.data(), o.N1, o.N1, phi,
                        true);
        }
      }
    }
    for (int c = 0; c < o.Nc; ++c)
      for (size_t i = 0; i < 2 * n0; ++i) {
        double &v = deltaNew[size_t(c) * 2 * n0 + i];
        v = o.domainRef[i % n0] != 0 ? v * o.coupling[c] : 0.0;
      }
    std::fill(aNew, aNew + 2 * n0, 0.0);
    if (doA) {
      const double jfac = static_cast<double>(2 * sgn(o.cfg.charge_sign)) * o.T;
      for (int f = 0; f < 2; ++f)
        for (size_t i = 0; i < n0; ++i)
          j[f * n0 + i] = o.domainRef[i] != 0 ? jacc[f * n0 + i] * jfac : 0.0;
      if (o.filter2D.empty())
        makeFilter2D(o.N0, o.h, o.filter2D, std::max(o.cfg.vector_potential_error, 0.0));
      solveVectorPotential(o.N0, o.h, o.numGrain, true, o.filter2D, j.data(), anew.data());
      const double l2 = o.lambda * o.lambda;
      for (size_t i = 0; i < 2 * n0; ++i) aNew[i] = anew[i] / l2;
    }
    o.bOutG = keepG;
    o.bOutGt = keepGt;
    return 0;
  } catch (const std::exception &e) {
    g_error = e.what();
    return -1;
  }
}

int oc_residual_gradient(void *hh, double *grad, double *gradA, double *residual) {
  try {
    Oracle &o = *static_cast<Oracle *>(hh);
    std::vector<double> g, ga;
    const double r = residualGradient(o, g, ga);
    std::copy(g.begin(), g.end(), grad);
    std::copy(ga.begin(), ga.end(), gradA);
    if (residual) *residual = r;
    return 0;
  } catch (const std::exception &e) {
    g_error = e.what();
    return -1;
  }
}

int oc_free_energy_gradient_a(void *hh, double *grad, double *gradA, double *free_energy) {
  try {
    Oracle &o = *static_cast<Oracle *>(hh);
    std::vector<double> g, ga;
    const double f = freeEnergyGradient(o, g, &ga);
    std::copy(g.begin(), g.end(), grad);
    std::copy(ga.begin(), ga.end(), gradA);
    if (free_energy) *free_energy = f;
    return 0;
  } catch (const std::exception &e) {
    g_error = e.what();
    return -1;
  }
}

long long oc_sweep_angles(void *hh, int a0, int a1) {
  Oracle &o = *static_cast<Oracle *>(hh);
  if (!o.boundaryInit) setBoundaryInitialCondition(o);
  std::vector<double> drot, es, gD, gJ, gF, tmp(2 * o.n0sq());
  long long steps = 0;
  for (int k = a0; k < a1; ++k) steps += integrateAngle(o, k, drot, es, gD, gJ, gF, tmp);
  return steps;
}

void oc_riccati_step(const double *d, const double *e, const double *i0,
                     double step, double *out) {
  const cplx r = riccatiStep(cplx(d[0], d[1]), cplx(e[0], e[1]),
                             cplx(i0[0], i0[1]), step);
  out[0] = r.real();
  out[1] = r.imag();
}
void oc_homogeneous(const double *d, const double *e, double *g, double *gt) {
  cplx a, b;
  homogeneous(cplx(d[0], d[1]), cplx(e[0], e[1]), a, b);
  g[0] = a.real();
  g[1] = a.imag();
  gt[0] = b.real();
  gt[1] = b.imag();
}
void oc_greens(const double *g, const double *gt, double *gg, double *fs) {
  const cplx a(g[0], g[1]), b(gt[0], gt[1]);
  const cplx x = greensG(a, b), y = greensDelta(a, b);
  gg[0] = x.real();
  gg[1] = x.imag();
  fs[0] = y.real();
  fs[1] = y.imag();
}
int oc_num_ozaki(double T, double cutoff) { return numOzakiEnergies(T, cutoff); }
int oc_ozaki(int n, double *p, double *r) {
  std::vector<double> pp, rr;
  ozakiPoles(n, pp, rr);
  std::copy(pp.begin(), pp.end(), p);
  std::copy(rr.begin(), rr.end(), r);
  return 0;
}
void oc_lattice(double w, double ppxi, int *n0, int *n1, double *gf) {
  lattice(w, ppxi, *n0, *n1, *gf);
}
double oc_bulk_gap(double T) { return bulkGap(T); }
double oc_bulk_free_energy(double T, double d) { return bulkFreeEnergy(T, d); }
double oc_yoshida(double T, double d, double c) { return yoshida(T, d, c); }
void *oc_accel_create(int kind, int n0, int num_fields, int rank, double step) {
  auto o = std::make_unique<Oracle>();
  o->cfg.accelerator = kind;
  o->cfg.rank = rank;
  o->cfg.step_size = step;
  o->N0 = n0;
  o->Nc = num_fields;
  return o.release();
}
void oc_accel_destroy(void *h) { delete static_cast<Oracle *>(h); }
int oc_accel_step(void *h, double *x, const double *g) {
  Oracle &o = *static_cast<Oracle *>(h);
  const size_t n = o.n0sq() * size_t(o.Nc);
  std::vector<double> X(x, x + n), G(g, g + n);
  switch (o.cfg.accelerator) {
  case OC_PICARD: picardStep(o, G, X); break;
  case OC_ANDERSON: andersonStep(o, G, X); break;
  case OC_CONGACC: congaccStep(o, G, X); break;
  case OC_BB: bbStep(o, G, X); break;
  case OC_POLYAK: polyakStep(o, G, X); break;
  case OC_BROYDEN: broydenStep(o, G, X); break;
  default: g_error = "unknown accelerator"; return -1;
  }
  std::copy(X.begin(), X.end(), x);
  return 0;
}
double oc_congacc_inertia(double dn, double mn, double dot, double min_cos) {
  return congaccInertia(dn, mn, dot, min_cos);
}
int oc_vector_potential(int n, double h, int ng, int enforce, const double *j,
                        double *out) {
  return oc_vector_potential_ex(n, h, ng, enforce, 0.0, 1, j, out);
}
int oc_vector_potential_ex(int n, double h, int ng, int enforce, double max_error,
                           int analytical, const double *j, double *out) {
  std::vector<double> filter;
  makeFilter2D(n, h, filter, max_error, analytical != 0);
  solveVectorPotential(n, h, ng, enforce != 0, filter, j, out);
  return 0;
}
int oc_poisson_kernel(int n, double h, double max_error, int analytical, double *out) {
  std::vector<double> filter;
  const int k = makeFilter2D(n, h, filter, max_error, analytical != 0);
  // the kernel at non-negative offsets, out[dy * n + dx]
  const int W = 2 * n - 1;
  for (int dy = 0; dy < n; ++dy)
    for (int dx = 0; dx < n; ++dx)
      out[size_t(dy) * n + dx] = filter[size_t(dy + n - 1) * W + size_t(dx + n - 1)];
  return k;
}

} // extern "C"
