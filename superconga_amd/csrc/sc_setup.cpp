// sc_setup.cpp -- host precompute of the iteration-invariant tables.
//
// Compiled with -ffp-contract=off: every discrete decision (domain labels,
// boundary crossings, reflection stencils) is taken on plainly rounded FP64
// arithmetic so that it is reproducible.

#include "sc_setup.hpp"
#include "sc_json.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <thread>

namespace scamd {

namespace {

constexpr int kMargin = 3;          // defines.h:25
constexpr double kMaxCoord = 0.5;   // defines.h:32
constexpr double kEulerGamma = 0.57721566490153286060651209008240243;

// Largest Ozaki pole when n+1 poles are used (Parameters.h:191-258); the
// reference's own table, used to choose the number of poles.
const double kPoleTable[] = {
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
constexpr int kPoleTableSize = sizeof(kPoleTable) / sizeof(double);

// Periodic wrap into [lo, hi) (utils.h:34-58).
double wrapRange(double v, double lo, double hi) {
  if (lo <= v && v < hi) return v;
  const double span = hi - lo;
  const double t = (v - lo) / span;
  return (t - std::floor(t)) * span + lo;
}
// Integer wrap through single precision with rounding (utils.h:70-78).
int wrapIndex(int v, int lo, int hi) {
  const float fv = static_cast<float>(v), flo = static_cast<float>(lo),
              fhi = static_cast<float>(hi);
  float r = fv;
  if (!(flo <= fv && fv < fhi)) {
    const float span = fhi - flo;
    const float t = (fv - flo) / span;
    r = (t - std::floor(t)) * span + flo;
  }
  return static_cast<int>(std::round(r));
}
inline void rot(double a, double x, double y, double &ox, double &oy) {
  const double c = std::cos(a), s = std::sin(a);
  ox = (x * c - y * s);
  oy = (x * s + y * c);
}
inline int signOf(double v) { return int(0.0 < v) - int(v < 0.0); }

double polarAngle(int k, int na) {
  const double u = M_PI * static_cast<double>(2 * k) / static_cast<double>(na) +
                   0.5 * M_PI;
  return wrapRange(u, 0.0, 2.0 * M_PI);
}
// Boole / Simpson 3/8 / Simpson / trapezoid weights (FermiSurface.h:218-249).
double angleWeight(int k, int na) {
  const int w = wrapIndex(k, 0, na);
  if (na % 4 == 0) return (w % 4 == 0) ? 28.0 / 45.0 : (w % 2 == 0 ? 24.0 / 45.0 : 64.0 / 45.0);
  if (na % 3 == 0) return (w % 3 == 0) ? 6.0 / 8.0 : 9.0 / 8.0;
  if (na % 2 == 0) return (w % 2 == 0) ? 2.0 / 3.0 : 4.0 / 3.0;
  return 1.0;
}
// FermiSurface.h:152-174 with a circle centred at the origin.
double fractionalAngleIndex(double px, double py, int na) {
  const double kx = wrapRange(px, -M_PI, M_PI);
  const double ky = wrapRange(py, -M_PI, M_PI);
  const double phi = wrapRange(std::atan2(ky, kx) - 0.5 * M_PI, 0.0, 2.0 * M_PI);
  return static_cast<double>(na) * phi / (2.0 * M_PI);
}

double basisValue(int sym, double px, double py) {
  switch (sym) {
  case SYM_S: return 1.0;
  case SYM_DX2Y2: return px * px - py * py;
  case SYM_DXY: return px * py;
  case SYM_G: return px * py * (px * px - py * py);
  }
  return 0.0;
}

// ---------------------------------------------------------------------------
// Geometry of one rotational frame.
// ---------------------------------------------------------------------------
struct Frame {
  std::vector<int8_t> label;
  std::vector<double> dist, nx, ny;
};

// 3x3 row-majority label of a disc (DiscGeometry.h:82-173).
int8_t discLabel(int X, int Y, double R, double ox, double oy) {
  const double R2 = R * R;
  bool row[3];
  for (int j = 0; j < 3; ++j) {
    const double dy = static_cast<double>(Y - 1 + j) - oy;
    int n = 0;
    for (int i = 0; i < 3; ++i) {
      const double dx = static_cast<double>(X - 1 + i) - ox;
      n += (dx * dx + dy * dy <= R2) ? 1 : 0;
    }
    row[j] = n > 1;
  }
  if (row[0] && row[1] && row[2]) return 1;
  if (!row[0] && row[1] && row[2]) return 2;
  if (row[0] && row[1] && !row[2]) return 3;
  return 0;
}
double toGrid(double f, int nSim, int nGrid) { // DiscGeometry.h:66-75
  return static_cast<double>(nSim - 1) * (f + kMaxCoord) +
         0.5 * static_cast<double>(nGrid - nSim);
}

void discFrame(const Shape &d, double angle, int N0, int N1, Frame &f) {
  double rx, ry;
  rot(angle, d.cx, d.cy, rx, ry);
  const double ox = toGrid(rx, N0, N1), oy = toGrid(ry, N0, N1);
  const double R = d.radius * static_cast<double>(N0 - 1);
  const size_t n = size_t(N1) * N1;
  f.label.assign(n, 0);
  f.dist.assign(n, 0.0);
  f.nx.assign(n, 0.0);
  f.ny.assign(n, 0.0);
  for (int ty = 0; ty < N1; ++ty)
    for (int tx = 0; tx < N1; ++tx) {
      const size_t i = size_t(ty) * N1 + tx;
      const int8_t l = discLabel(tx, ty, R, ox, oy);
      f.label[i] = l;
      if (l > 1) {
        const double x = static_cast<double>(tx) - ox;
        const double y = static_cast<double>(ty) - oy;
        const double dd = std::sqrt(R * R - x * x) - std::fabs(y);
        f.dist[i] = dd;
        f.nx[i] = -x / R;
        f.ny[i] = (y <= 0.0) ? -(y - dd) / R : -(y + dd) / R;
      }
    }
}

// PolygonGeometry.h:101-190, GeometryComponent.h:112-147, :233-251.
void polygonFrame(const Shape &p, double angle, int N0, int N1, Frame &f) {
  const int nv = static_cast<int>(p.vx.size());
  const double sc = static_cast<double>(N0 - 1);
  const double ca = std::cos(angle), sa = std::sin(angle);
  std::vector<double> vx(nv), vy(nv), ex(nv), ey(nv), nxv(nv), nyv(nv);
  for (int i = 0; i < nv; ++i) {
    vx[i] = (ca * p.vx[i] - sa * p.vy[i]) * sc;
    vy[i] = (sa * p.vx[i] + ca * p.vy[i]) * sc;
  }
  for (int i = 0; i < nv; ++i) {
    const int j = (i + 1 == nv) ? 0 : i + 1;
    ex[i] = vx[j] - vx[i];
    ey[i] = vy[j] - vy[i];
    const double inv = 1.0 / std::sqrt(ey[i] * ey[i] + ex[i] * ex[i]);
    nxv[i] = -ey[i] * inv;
    nyv[i] = ex[i] * inv;
  }
  const size_t n = size_t(N1) * N1;
  std::vector<int8_t> in(n, 0);
  f.dist.assign(n, 0.0);
  f.nx.assign(n, 0.0);
  f.ny.assign(n, 0.0);
  const double half = static_cast<double>(N1 - 1) * 0.5;
  for (int ty = 0; ty < N1; ++ty)
    for (int tx = 0; tx < N1; ++tx) {
      const double X = static_cast<double>(tx) - half;
      const double Y = static_cast<double>(ty) - half;
      int cnt = 0;
      double best = 100000.0, bnx = 0.0, bny = 0.0;
      for (int i = 0; i < nv; ++i) {
        if (ex[i] * (Y - vy[i]) - ey[i] * (X - vx[i]) >= 0.0) ++cnt;
        double d = 10000.0;
        if (std::fabs(ex[i]) > 0.00001) {
          const double k = ey[i] / ex[i];
          d = std::fabs(k * X + (vy[i] - k * vx[i]) - Y);
        }
        if (d < best) {
          best = d;
          bnx = nxv[i];
          bny = nyv[i];
        }
      }
      const size_t i = size_t(ty) * N1 + tx;
      in[i] = (cnt == nv) ? 1 : 0;
      f.dist[i] = best;
      f.nx[i] = bnx;
      f.ny[i] = bny;
    }
  f.label.assign(n, 0);
  for (int ty = 1; ty < N1 - 1; ++ty)
    for (int tx = 0; tx < N1; ++tx) {
      const size_t i = size_t(ty) * N1 + tx;
      if (in[i] != 1) continue;
      const bool below = in[i - N1] == 0, above = in[i + N1] == 0;
      f.label[i] = (below && above) ? 0 : (below ? 2 : (above ? 3 : 1));
    }
  for (size_t i = 0; i < n; ++i)
    if (f.label[i] <= 1) {
      f.dist[i] *= 0.0;
      f.nx[i] *= 0.0;
      f.ny[i] *= 0.0;
    }
}

// GeometryGroup::create (GeometryGroup.h:216-319,487-522).
void groupFrame(const std::vector<Shape> &shapes, double angle, int N0, int N1,
                Frame &g) {
  static const int8_t opAdd[16] = {0, 1, 2, 3, 1, 1, 1, 1,
                                   2, 1, 2, 1, 3, 1, 1, 3};
  const size_t n = size_t(N1) * N1;
  g.label.assign(n, 0);
  g.dist.assign(n, 0.0);
  g.nx.assign(n, 0.0);
  g.ny.assign(n, 0.0);
  Frame c;
  for (const Shape &s : shapes) {
    if (s.disc)
      discFrame(s, angle, N0, N1, c);
    else
      polygonFrame(s, angle, N0, N1, c);
    if (s.add) {
      for (size_t i = 0; i < n; ++i) {
        const int8_t l = opAdd[g.label[i] * 4 + c.label[i]];
        double ax = 0.0, ay = 0.0;
        if (l == 2 || l == 3) {
          ax = g.nx[i] + c.nx[i];
          ay = g.ny[i] + c.ny[i];
          const double inv = 1.0 / std::sqrt(ax * ax + ay * ay);
          ax *= inv;
          ay *= inv;
        }
        g.nx[i] = ax;
        g.ny[i] = ay;
        g.label[i] = l;
        g.dist[i] = std::fmax(g.dist[i], c.dist[i]);
      }
    } else {
      std::vector<int8_t> out(n, 0);
      for (int ty = 1; ty < N1 - 1; ++ty)
        for (int tx = 0; tx < N1; ++tx) {
          const size_t i = size_t(ty) * N1 + tx;
          int8_t l = g.label[i];
          const size_t src = (c.label[i + N1] == 2) ? i + N1 : i - N1;
          if (c.label[i] > 0) {
            l = 0;
          } else if ((c.label[i + N1] == 2 || c.label[i - N1] == 3) && l == 1) {
            l = (c.label[i + N1] == 2) ? 3 : 2;
            double ax = g.nx[i] - c.nx[src], ay = g.ny[i] - c.ny[src];
            const double inv = 1.0 / std::sqrt(ax * ax + ay * ay);
            g.nx[i] = ax * inv;
            g.ny[i] = ay * inv;
            g.dist[i] = 1.0 - c.dist[src];
          }
          out[i] = l;
        }
      g.label.swap(out);
    }
  }
}

// GeometryGroup::createDomain on the reference N0 grid.
std::vector<uint8_t> referenceDomain(const std::vector<Shape> &shapes, int N0) {
  const size_t n = size_t(N0) * N0;
  std::vector<uint8_t> dom(n, 0);
  std::vector<uint8_t> loc(n);
  for (const Shape &s : shapes) {
    if (s.disc) {
      const double ox = toGrid(s.cx, N0, N0), oy = toGrid(s.cy, N0, N0);
      const double R = s.radius * static_cast<double>(N0 - 1);
      for (int ty = 0; ty < N0; ++ty)
        for (int tx = 0; tx < N0; ++tx)
          loc[size_t(ty) * N0 + tx] = discLabel(tx, ty, R, ox, oy) > 0;
    } else {
      const int nv = static_cast<int>(s.vx.size());
      for (int ty = 0; ty < N0; ++ty)
        for (int tx = 0; tx < N0; ++tx) {
          const double X = static_cast<double>(tx) / static_cast<double>(N0 - 1) - 0.5;
          const double Y = static_cast<double>(ty) / static_cast<double>(N0 - 1) - 0.5;
          int cnt = 0;
          for (int i = 0; i < nv; ++i) {
            const int j = (i == nv - 1) ? 0 : i + 1;
            const double lx = s.vx[j] - s.vx[i], ly = s.vy[j] - s.vy[i];
            if (lx * (Y - s.vy[i]) - ly * (X - s.vx[i]) >= 0.0) ++cnt;
          }
          loc[size_t(ty) * N0 + tx] = cnt == nv;
        }
    }
    for (size_t i = 0; i < n; ++i)
      dom[i] = s.add ? (dom[i] | loc[i]) : ((dom[i] && !loc[i]) ? 1 : 0);
  }
  return dom;
}

void regularPolygon(int n, double radius, double angle, double cx, double cy,
                    std::vector<double> &vx, std::vector<double> &vy) {
  // Polygon.h:536-561
  vx.resize(n);
  vy.resize(n);
  const double step = 2.0 * M_PI / static_cast<double>(n);
  const double first = step * 0.5 + angle;
  for (int i = 0; i < n; ++i) {
    const double a = step * static_cast<double>(i) + first;
    vx[i] = radius * std::cos(a) + cx;
    vy[i] = radius * std::sin(a) + cy;
  }
}

// Label>0-neighbourhood extension lists (domain_ops.h:11-57).
void extensionList(const int8_t *labels, int N, int base, ExtensionList &out) {
  for (int y = 0; y < N; ++y)
    for (int x = 0; x < N; ++x) {
      if (labels[size_t(y) * N + x] != 0) continue;
      int best = 1000;
      std::vector<int> src;
      for (int ky = std::max(y - 2, 0); ky <= std::min(y + 2, N - 1); ++ky)
        for (int kx = std::max(x - 2, 0); kx <= std::min(x + 2, N - 1); ++kx) {
          if (labels[size_t(ky) * N + kx] <= 0) continue;
          const int d2 = (x - kx) * (x - kx) + (y - ky) * (y - ky);
          if (d2 < best) {
            best = d2;
            src.assign(1, base + ky * N + kx);
          } else if (d2 == best) {
            src.push_back(base + ky * N + kx);
          }
        }
      if (src.empty()) continue;
      out.target.push_back(base + y * N + x);
      out.source.insert(out.source.end(), src.begin(), src.end());
      out.offset.push_back(static_cast<int>(out.source.size()));
    }
}

template <class F> void parallelFor(int n, F &&f) {
  const int nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  std::atomic<int> next{0};
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&] {
      for (int i; (i = next.fetch_add(1)) < n;) f(i);
    });
  for (auto &th : pool) th.join();
}

} // namespace

// ---------------------------------------------------------------------------
// Public helpers
// ---------------------------------------------------------------------------
int numOzakiEnergies(double temperature, double cutoff) {
  // Parameters.h:380-420
  const double target = 2.0 * M_PI * (cutoff / temperature);
  const double *it = std::upper_bound(kPoleTable, kPoleTable + kPoleTableSize, target);
  if (it != kPoleTable + kPoleTableSize) return static_cast<int>(it - kPoleTable) + 1;
  int n = kPoleTableSize - 1;
  for (;;) {
    ++n;
    const double x = n;
    if (x * (2.5465 * x + 1.2744) + 0.4651 >= target) return n;
  }
}

// The 2N x 2N Ozaki problem A x = w B x (B = diag(2i+1), A = -1/2 on the
// first off-diagonals) is symmetric tridiagonal after the scaling
// B^-1/2 A B^-1/2.  It is diagonalized with the implicit QL iteration with
// Wilkinson shifts (the EISPACK tql2 scheme, Bowdler, Martin, Reinsch &
// Wilkinson 1968), tracking only the first row of the eigenvector matrix:
// the residue needs the first component of each B-normalized eigenvector.
void ozakiPoles(int numPoles, std::vector<double> &poles,
                std::vector<double> &residues) {
  const int n = 2 * numPoles;
  std::vector<double> d(n, 0.0), e(n, 0.0), z(n, 0.0);
  for (int i = 0; i + 1 < n; ++i)
    e[i] = -0.5 / std::sqrt((2.0 * i + 1.0) * (2.0 * i + 3.0));
  z[0] = 1.0;
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < n; ++l) {
    for (int iter = 0; iter < 200; ++iter) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m == l) break;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
      double s = 1.0, c = 1.0, p = 0.0;
      int i = m - 1;
      for (; i >= l; --i) {
        double f = s * e[i];
        const double b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        f = z[i + 1];
        z[i + 1] = s * z[i] + c * f;
        z[i] = c * z[i] - s * f;
      }
      if (r == 0.0 && i >= l) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return d[a] < d[b]; });
  poles.resize(numPoles);
  residues.resize(numPoles);
  for (int i = 0; i < numPoles; ++i) {
    const double w = d[ord[i]];
    const double pole = -1.0 / w;
    const double half = 0.5 * z[ord[i]] * pole;
    poles[i] = pole;
    residues[i] = half * half;
  }
}

// OrderParameterBulk.h:88-220 (s-wave bulk gap; Matsubara cutoff 100).
double bulkGapSWave(double T) {
  if (T >= 1.0) return 0.0;
  const double e0 = 0.5 * std::exp(-kEulerGamma);
  if (T < std::numeric_limits<double>::epsilon()) return e0;
  const int n = static_cast<int>(100.0 / T - 0.5) + 1;
  double inv = 0.0;
  for (int k = n - 1; k >= 0; --k) inv += 1.0 / (T * (k + 0.5));
  const double g = 1.0 / (std::log(T) / T + inv);
  auto residual = [&](double x) {
    double s = 0.0;
    for (int k = n - 1; k >= 0; --k) s += 1.0 / std::hypot(x, T * (k + 0.5));
    return x * (s * g - 1.0);
  };
  double x = e0 * std::sqrt(1.0 - T * T * T);
  double xs[2] = {x, std::sqrt(x)}, rs[2] = {residual(xs[0]), residual(xs[1])};
  const double crit = 10.0 * std::numeric_limits<double>::epsilon();
  for (int it = 0; it < 10; ++it) {
    const double w = 1.0 / (1.0 - (xs[0] / xs[1]) / (rs[0] / rs[1]));
    x = (1.0 - w) * xs[0] + w * xs[1];
    const double r = residual(x);
    const double rel =
        x < 10.0 * std::numeric_limits<double>::epsilon() ? r : std::fabs(r / x);
    if (rel < crit) break;
    const int worst = std::fabs(rs[0]) > std::fabs(rs[1]) ? 0 : 1;
    xs[worst] = x;
    rs[worst] = r;
  }
  return x;
}

void latticeSizes(double width, double ppxi, int &n0, int &n1, double &gf) {
  // Parameters.h:422-458
  const double pts = width * ppxi;
  n0 = static_cast<int>(std::ceil(pts)) + 2 * kMargin;
  gf = pts / static_cast<double>(n0);
  n1 = n0 + 2 * static_cast<int>(
                    std::ceil(0.5 * (std::sqrt(2.0) - 1.0) * static_cast<double>(n0)));
}

// Eigen-decomposition of a real symmetric n x n matrix (row-major in a,
// overwritten): Householder reduction to tridiagonal form (Householder
// 1958), the orthogonal factor formed explicitly, then the implicit QL
// iteration with Wilkinson shifts (the EISPACK tql2 scheme) with every plane
// rotation applied to that factor.  The factor is kept transposed, so that a
// rotation touches two contiguous rows: on return row k of vecs is the unit
// eigenvector of eigenvalue vals[k] (in no particular order).
void symmetricEigen(int n, std::vector<double> &a, std::vector<double> &vals,
                    std::vector<double> &vecs) {
  const size_t N = size_t(n);
  std::vector<double> d(N, 0.0), e(N, 0.0);
  std::vector<double> house; // Householder vectors, v_k over rows k+1..n-1
  std::vector<double> beta(N, 0.0);
  std::vector<size_t> vOff(N, 0);
  std::vector<double> p(N), w(N);
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    vOff[size_t(k)] = house.size();
    double xx = 0.0;
    for (int r = k + 1; r < n; ++r) xx += a[size_t(r) * N + k] * a[size_t(r) * N + k];
    const double x0 = a[size_t(k + 1) * N + k];
    const double norm = std::sqrt(xx);
    d[size_t(k)] = a[size_t(k) * N + k];
    if (norm == 0.0) {
      e[size_t(k)] = 0.0;
      house.insert(house.end(), size_t(m), 0.0);
      continue;
    }
    const double alpha = x0 >= 0.0 ? -norm : norm;
    e[size_t(k)] = alpha;
    // v = x - alpha e_1, beta = 2 / |v|^2
    house.push_back(x0 - alpha);
    for (int r = k + 2; r < n; ++r) house.push_back(a[size_t(r) * N + k]);
    const double *v = house.data() + vOff[size_t(k)];
    double vv = 0.0;
    for (int i = 0; i < m; ++i) vv += v[i] * v[i];
    const double b = 2.0 / vv;
    beta[size_t(k)] = b;
    // trailing block B <- H B H: p = b B v, w = p - (b/2)(p.v) v,
    // B <- B - v w^T - w v^T
    for (int i = 0; i < m; ++i) {
      const double *row = &a[size_t(k + 1 + i) * N + size_t(k + 1)];
      double s = 0.0;
      for (int j = 0; j < m; ++j) s += row[j] * v[j];
      p[size_t(i)] = b * s;
    }
    double pv = 0.0;
    for (int i = 0; i < m; ++i) pv += p[size_t(i)] * v[i];
    const double kk = 0.5 * b * pv;
    for (int i = 0; i < m; ++i) w[size_t(i)] = p[size_t(i)] - kk * v[i];
    for (int i = 0; i < m; ++i) {
      double *row = &a[size_t(k + 1 + i) * N + size_t(k + 1)];
      const double vi = v[i], wi = w[size_t(i)];
      for (int j = 0; j < m; ++j) row[j] -= vi * w[size_t(j)] + wi * v[j];
    }
  }
  if (n >= 2) {
    d[N - 2] = a[(N - 2) * N + (N - 2)];
    e[N - 2] = a[(N - 1) * N + (N - 2)];
  }
  d[N - 1] = a[(N - 1) * N + (N - 1)];
  e[N - 1] = 0.0;
  // the transposed orthogonal factor H_{n-3} ... H_0 (each H symmetric)
  vecs.assign(N * N, 0.0);
  for (size_t i = 0; i < N; ++i) vecs[i * N + i] = 1.0;
  std::vector<double> u(N);
  for (int k = 0; k + 2 < n; ++k) {
    const double b = beta[size_t(k)];
    if (b == 0.0) continue;
    const double *v = house.data() + vOff[size_t(k)];
    const int m = n - k - 1;
    std::fill(u.begin(), u.end(), 0.0);
    for (int i = 0; i < m; ++i) {
      const double *row = &vecs[size_t(k + 1 + i) * N];
      const double vi = v[i];
      for (size_t j = 0; j < N; ++j) u[j] += vi * row[j];
    }
    for (int i = 0; i < m; ++i) {
      double *row = &vecs[size_t(k + 1 + i) * N];
      const double f = b * v[i];
      for (size_t j = 0; j < N; ++j) row[j] -= f * u[j];
    }
  }
  // implicit QL on (d, e), e[i] coupling i and i + 1
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < n; ++l) {
    for (int iter = 0; iter < 300; ++iter) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = std::fabs(d[size_t(m)]) + std::fabs(d[size_t(m + 1)]);
        if (std::fabs(e[size_t(m)]) <= eps * dd) break;
      }
      if (m == l) break;
      if (iter == 299) throw std::runtime_error("eigen-decomposition did not converge");
      double g = (d[size_t(l + 1)] - d[size_t(l)]) / (2.0 * e[size_t(l)]);
      double r = std::hypot(g, 1.0);
      g = d[size_t(m)] - d[size_t(l)] + e[size_t(l)] / (g + (g >= 0.0 ? r : -r));
      double s = 1.0, c = 1.0, pp = 0.0;
      int i = m - 1;
      for (; i >= l; --i) {
        double f = s * e[size_t(i)];
        const double bb = c * e[size_t(i)];
        r = std::hypot(f, g);
        e[size_t(i + 1)] = r;
        if (r == 0.0) {
          d[size_t(i + 1)] -= pp;
          e[size_t(m)] = 0.0;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[size_t(i + 1)] - pp;
        r = (d[size_t(i)] - g) * s + 2.0 * c * bb;
        pp = s * r;
        d[size_t(i + 1)] = g + pp;
        g = c * r - bb;
        double *ri = &vecs[size_t(i) * N], *rj = &vecs[size_t(i + 1) * N];
        for (size_t q = 0; q < N; ++q) {
          const double t = rj[q];
          rj[q] = s * ri[q] + c * t;
          ri[q] = c * ri[q] - s * t;
        }
      }
      if (r == 0.0 && i >= l) continue;
      d[size_t(l)] -= pp;
      e[size_t(l)] = g;
      e[size_t(m)] = 0.0;
    }
  }
  vals = d;
}

int poissonKernel(int n, double h, std::vector<double> &g, double maxError) {
  // ComputeVectorPotential.h:212-294: analytic matrix elements, symmetrized,
  // scaled by 1/4pi, mean moved into the constant part, then G = Q + C.
  auto alpha = [](int a, int b) {
    return static_cast<double>(a * b) *
           (2.0 * std::log(std::hypot(double(a), double(b))) - std::log(2.0));
  };
  auto beta = [](int a, int b) {
    return 0.5 * static_cast<double>((2 - a + b) * (a + b) - 2) *
           std::atan(static_cast<double>(a) / static_cast<double>(b));
  };
  std::vector<double> q(size_t(n) * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      double acc = 0.0;
      for (int jj : {-j, j})
        for (int ii : {-i, i}) {
          const int a = 1 + 2 * ii, b = 1 + 2 * jj;
          acc += alpha(a, b) + beta(a, b) + beta(b, a);
        }
      q[size_t(j) * n + i] = acc * 0.25; // element (i, j)
    }
  const double k = 1.0 / (4.0 * M_PI);
  g.assign(size_t(n) * n, 0.0);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i)
      g[size_t(j) * n + i] = 0.5 * (q[size_t(j) * n + i] + q[size_t(i) * n + j]) * k;
  double c = (2.0 * std::log(h) - std::log(2.0) - 3.0) * k;
  double mean = 0.0;
  for (int j = 0; j < n; ++j) {
    double col = 0.0;
    for (int i = 0; i < n; ++i) col += g[size_t(j) * n + i];
    mean += col / n;
  }
  mean /= n;
  c += mean;
  for (double &v : g) v -= mean;
  if (maxError <= 0.0) {
    // g[dy*n + dx] = G(|dx|, |dy|); the symmetrized matrix is symmetric.
    for (double &v : g) v += c;
    return 0;
  }
  // vector_potential_error > 0 (ComputeVectorPotential.h:319-395): Q is
  // replaced by its eigen-expansion truncated to the eigenvalues of largest
  // magnitude, added until the relative Frobenius error |Q - Q_K| / |Q|
  // falls below maxError; the constant C is exact (a rank-one filter of
  // ones).  The reference applies the expansion as K separable filter
  // pairs; its sum is the same kernel, which the FFT convolution applies at
  // the cost of the exact one.
  std::vector<double> vals, vecs;
  symmetricEigen(n, g, vals, vecs);
  std::vector<int> order(static_cast<size_t>(n));
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return std::fabs(vals[size_t(a)]) > std::fabs(vals[size_t(b)]);
  });
  double total = 0.0;
  for (double v : vals) total += v * v;
  std::vector<int> kept;
  double partial = 0.0;
  for (int k : order) {
    kept.push_back(k);
    partial += vals[size_t(k)] * vals[size_t(k)];
    const double err2 = 1.0 - partial / total;
    if (std::sqrt(std::max(err2, 0.0)) < maxError) break;
  }
  g.assign(size_t(n) * n, std::fabs(c) > 0.0 ? c : 0.0);
  for (int k : kept) {
    const double lam = vals[size_t(k)];
    const double *v = &vecs[size_t(k) * n];
    for (int j = 0; j < n; ++j) {
      const double f = lam * v[j];
      double *row = &g[size_t(j) * n];
      for (int i = 0; i < n; ++i) row[i] += f * v[i];
    }
  }
  return static_cast<int>(kept.size()) + (std::fabs(c) > 0.0 ? 1 : 0);
}

// ---------------------------------------------------------------------------
// Configuration (src/simulation.cu:64-463, geometry_utils.h:41-325)
// ---------------------------------------------------------------------------
Config parseConfig(const std::string &text) {
  const json::Value root = json::parse(text);
  Config c;
  const json::Value &ph = root["physics"];
  c.temperature = ph["temperature"].asDouble(0.5);
  c.fluxQuanta = ph["external_flux_quanta"].asDouble(0.0);
  c.lambda = ph["penetration_depth"].asDouble(-1.0);
  c.crystalRotation = 2.0 * M_PI * ph["crystal_axes_rotation"].asDouble(0.0);
  const std::string gauge = ph["gauge"].asString("symmetric");
  if (gauge == "landau") c.gauge = GAUGE_LANDAU;
  else if (gauge == "symmetric") c.gauge = GAUGE_SYMMETRIC;
  else if (gauge == "solenoid") c.gauge = GAUGE_SOLENOID;
  else throw std::runtime_error("unknown gauge '" + gauge + "'");
  c.chargeSign = signOf(ph["charge_sign"].asDouble(-1));
  const json::Value &op = ph["order_parameter"];
  const char *names[4] = {"dx2-y2", "dxy", "g", "s"}; // simulation.cu:201
  const int syms[4] = {SYM_DX2Y2, SYM_DXY, SYM_G, SYM_S};
  for (int i = 0; i < 4; ++i) {
    if (!op.has(names[i])) continue;
    const json::Value &p = op[names[i]];
    const int k = c.numComponents++;
    c.symmetry[k] = syms[i];
    c.tc[k] = p["critical_temperature"].asDouble(1.0);
    c.noise[k] = p["initial_noise_stddev"].asDouble(0.0);
    c.phaseShift[k] = 2.0 * M_PI * p["initial_phase_shift"].asDouble(0.0);
    for (size_t v = 0; v < p["vortices"].size(); ++v) {
      const json::Value &vv = p["vortices"][v];
      c.vortices[k].push_back({vv["center_x"].asDouble(), vv["center_y"].asDouble(),
                               vv["winding_number"].asDouble()});
    }
  }
  if (c.numComponents == 0) throw std::runtime_error("no order-parameter component set");

  const json::Value &nu = root["numerics"];
  c.convergence = nu["convergence_criterion"].asDouble(1e-6);
  c.cutoff = nu["energy_cutoff"].asDouble(10.0);
  const std::string norm = nu["norm"].asString("l2");
  if (norm == "l1") c.norm = NORM_L1;
  else if (norm == "l2") c.norm = NORM_L2;
  else if (norm == "linf") c.norm = NORM_LINF;
  else throw std::runtime_error("unknown norm '" + norm + "'");
  c.numAngles = nu["num_fermi_momenta"].asInt(64);
  c.ppxi = nu["points_per_coherence_length"].asDouble(5.0);
  c.vpError = nu["vector_potential_error"].asDouble(0.0);
  c.burnin = nu["num_iterations_burnin"].asInt(0);
  c.itMin = nu["num_iterations_min"].asInt(0);
  c.itMax = nu["num_iterations_max"].asInt(10000);
  if (c.numAngles < 1) throw std::runtime_error("num_fermi_momenta must be positive");
  if (!(c.temperature > 0.0)) throw std::runtime_error("temperature must be positive");
  if (!(c.ppxi > 0.0)) throw std::runtime_error("points_per_coherence_length must be positive");

  const json::Value &ac = root["accelerator"];
  const std::string an = ac["name"].asString("congacc");
  if (an == "picard") c.accel = ACC_PICARD;
  else if (an == "anderson") c.accel = ACC_ANDERSON;
  else if (an == "congacc") c.accel = ACC_CONGACC;
  else if (an == "barzilai-borwein" || an == "bb") c.accel = ACC_BB;
  else if (an == "polyak") c.accel = ACC_POLYAK;
  else if (an == "broyden") c.accel = ACC_BROYDEN;
  else throw std::runtime_error("unknown accelerator '" + an + "'");
  c.stepSize = ac["step_size"].asDouble(-1);
  c.rank = ac["rank"].asInt(-1);
  c.blockScaling = ac["block_scaling"].asBool(false);
  c.vpScreening = ac["vector_potential_screening"].asDouble(0.0);
  if (c.vpScreening < 0.0)
    throw std::runtime_error("accelerator.vector_potential_screening must be >= 0");
  c.drag = ac["drag"].asDouble(-1);
  c.stepFactor = ac["step_size_factor"].asDouble(-1);
  c.cosMin = ac["cos_similarity_min"].asDouble(-1);
  c.stepMax = ac["step_size_max"].asDouble(-1);
  c.stepMin = ac["step_size_min"].asDouble(-1);

  const json::Value &misc = root["misc"];
  c.loadPath = misc["load_path"].asString("");
  c.verbose = misc["verbose"].asBool(false);
  c.seed = static_cast<unsigned long long>(misc["seed"].asDouble(0));
  const json::Value &comp = root["compute"];
  c.computeCurrent = comp["current"].asBool(true);
  c.computeFreeEnergy = comp["free_energy"].asBool(true);

  // Geometry.  Internal-frame components (used to restate the reference's
  // own tests, which build DiscGeometry/PolygonGeometry directly) or the
  // user-facing "geometry" list scaled by its bounding box.
  struct Raw {
    int kind; // 0 disc, 1 regular polygon, 2 polygon
    bool add;
    double p[5];
    std::vector<double> vx, vy;
  };
  std::vector<Raw> raw;
  const bool internal = root.has("geometry_internal");
  const json::Value &glist =
      internal ? root["geometry_internal"]["components"] : root["geometry"];
  for (size_t i = 0; i < glist.size(); ++i) {
    const json::Value &g = glist[i];
    if (g.kind != json::Value::Object || g.obj.empty())
      throw std::runtime_error("bad geometry entry");
    const std::string kind = g.obj[0].first;
    const json::Value &p = g.obj[0].second;
    Raw r{};
    r.add = p["add"].asBool(true);
    if (kind == "disc") {
      r.kind = 0;
      r.p[0] = p["center_x"].asDouble(0);
      r.p[1] = p["center_y"].asDouble(0);
      r.p[2] = p["radius"].asDouble(internal ? kMaxCoord : 1.0);
    } else if (kind == "regular_polygon") {
      r.kind = 1;
      r.p[0] = p["center_x"].asDouble(0);
      r.p[1] = p["center_y"].asDouble(0);
      r.p[2] = p["num_edges"].asInt(4);
      r.p[3] = p["rotation"].asDouble(0);
      r.p[4] = p["side_length"].asDouble(-1);
    } else if (kind == "polygon") {
      r.kind = 2;
      for (size_t v = 0; v < p["vertices_x"].size(); ++v) {
        r.vx.push_back(p["vertices_x"][v].asDouble());
        r.vy.push_back(p["vertices_y"][v].asDouble());
      }
      if (r.vx.size() != p["vertices_y"].size() || r.vx.size() < 3)
        throw std::runtime_error("polygon needs matching vertices_x/vertices_y");
    } else {
      throw std::runtime_error("unknown geometry type '" + kind + "'");
    }
    raw.push_back(r);
  }
  if (raw.empty()) throw std::runtime_error("no geometry given");

  double offX = 0, offY = 0, width;
  if (internal) {
    width = root["geometry_internal"]["grain_width"].asDouble(20.0);
  } else {
    double xMax = std::numeric_limits<double>::min(),
           xMin = std::numeric_limits<double>::max(),
           yMax = std::numeric_limits<double>::min(),
           yMin = std::numeric_limits<double>::max();
    auto extend = [&](double x, double y) {
      xMax = std::max(xMax, x);
      xMin = std::min(xMin, x);
      yMax = std::max(yMax, y);
      yMin = std::min(yMin, y);
    };
    for (const Raw &r : raw) {
      if (!r.add) continue;
      if (r.kind == 0) {
        extend(r.p[0] + r.p[2], r.p[1] + r.p[2]);
        extend(r.p[0] - r.p[2], r.p[1] - r.p[2]);
      } else if (r.kind == 2) {
        for (size_t v = 0; v < r.vx.size(); ++v) extend(r.vx[v], r.vy[v]);
      } else {
        const int ne = static_cast<int>(r.p[2]);
        const double step = 2.0 * M_PI / static_cast<double>(ne);
        const double rad = r.p[4] / std::hypot(std::cos(step) - 1.0, std::sin(step));
        const double first = 0.5 * step + 2.0 * M_PI * r.p[3];
        for (int v = 0; v < ne; ++v) {
          const double a = first + step * static_cast<double>(v);
          extend(r.p[0] + rad * std::cos(a), r.p[1] + rad * std::sin(a));
        }
      }
    }
    const double xe = xMax - xMin, ye = yMax - yMin;
    offX = xMin + 0.5 * xe;
    offY = yMin + 0.5 * ye;
    width = std::max(xe, ye);
  }
  c.grainWidth = width;
  c.geomOffX = offX;
  c.geomOffY = offY;
  int n0, n1;
  double gf;
  latticeSizes(width, c.ppxi, n0, n1, gf);
  for (const Raw &r : raw) {
    Shape s;
    s.add = r.add;
    if (r.kind == 0) {
      const double cx = internal ? r.p[0] : (r.p[0] - offX) / width;
      const double cy = internal ? r.p[1] : (r.p[1] - offY) / width;
      const double rr = internal ? r.p[2] : r.p[2] / width;
      s.disc = true;
      s.radius = rr * gf;
      s.cx = cx * gf;
      s.cy = cy * gf;
    } else if (r.kind == 2) {
      s.disc = false;
      for (size_t v = 0; v < r.vx.size(); ++v) {
        s.vx.push_back((internal ? r.vx[v] : (r.vx[v] - offX) / width) * gf);
        s.vy.push_back((internal ? r.vy[v] : (r.vy[v] - offY) / width) * gf);
      }
    } else {
      const int ne = static_cast<int>(r.p[2]);
      const double cx = internal ? r.p[0] : (r.p[0] - offX) / width;
      const double cy = internal ? r.p[1] : (r.p[1] - offY) / width;
      const double side = internal ? r.p[4] : r.p[4] / width;
      const double rotation = internal ? r.p[3] : 2.0 * M_PI * r.p[3];
      const double step = 2.0 * M_PI / static_cast<double>(ne);
      const double rad = side < 0.0 ? kMaxCoord
                                    : side / std::hypot(std::cos(step) - 1.0, std::sin(step));
      s.disc = false;
      regularPolygon(ne, rad * gf, rotation, cx * gf, cy * gf, s.vx, s.vy);
    }
    c.shapes.push_back(s);
  }
  return c;
}

// ---------------------------------------------------------------------------
// Setup
// ---------------------------------------------------------------------------
Spectroscopy parseSpectroscopy(const std::string &text) {
  const json::Value root = json::parse(text);
  Spectroscopy sp;
  const json::Value &sc = root["spectroscopy"];
  sp.eMin = sc["energy_min"].asDouble(0.0);
  sp.eMax = sc["energy_max"].asDouble(0.0);
  sp.broadening = sc["energy_broadening"].asDouble(0.0);
  sp.numEnergies = sc["num_energies"].asInt(0);
  const json::Value &nu = root["numerics"];
  sp.convergence = nu["convergence_criterion"].asDouble(1e-4);
  const std::string norm = nu["norm"].asString("l2");
  if (norm == "l1") sp.norm = NORM_L1;
  else if (norm == "l2") sp.norm = NORM_L2;
  else if (norm == "linf") sp.norm = NORM_LINF;
  else throw std::runtime_error("unknown norm '" + norm + "'");
  sp.perBlock = nu["num_energies_per_block"].asInt(-1);
  sp.numAngles = nu["num_fermi_momenta"].asInt(0);
  sp.burnin = nu["num_iterations_burnin"].asInt(-1);
  sp.itMin = nu["num_iterations_min"].asInt(0);
  sp.itMax = nu["num_iterations_max"].asInt(-1);
  if (sp.numEnergies < 1) throw std::runtime_error("num_energies must be positive");
  if (sp.numAngles < 1) throw std::runtime_error("num_fermi_momenta must be positive");
  if (!(sp.broadening > 0.0)) throw std::runtime_error("energy_broadening must be positive");
  if (sp.perBlock == 0) throw std::runtime_error("num_energies_per_block must not be 0");
  // IntegrationIteratorKeldysh::initIterators (:152-180)
  sp.perBlock = sp.perBlock < 1 ? sp.numEnergies : std::min(sp.perBlock, sp.numEnergies);
  return sp;
}

void setKeldyshBlock(Setup &s, int start) {
  // IntegrationIteratorKeldysh::initEnergies (:115-150): E_i = min + i step
  const Spectroscopy &sp = s.spec;
  const int n = sp.numEnergies;
  const double step = n > 1 ? (sp.eMax - sp.eMin) / static_cast<double>(n - 1) : 0.0;
  s.Ne = std::min(sp.perBlock, n - start);
  s.energyRe.resize(size_t(s.Ne));
  s.energy.assign(size_t(s.Ne), sp.broadening);
  s.residue.assign(size_t(s.Ne), 1.0);
  for (int e = 0; e < s.Ne; ++e) s.energyRe[size_t(e)] = sp.eMin + step * static_cast<double>(start + e);
  s.trajSteps = 2LL * s.Ne * s.insideSites;
}

std::unique_ptr<Setup> buildSpectroscopySetup(Config cfg, const Spectroscopy &sp) {
  cfg.numAngles = sp.numAngles;
  cfg.convergence = sp.convergence;
  cfg.norm = sp.norm;
  cfg.burnin = sp.burnin;
  cfg.itMin = sp.itMin;
  cfg.itMax = sp.itMax;
  cfg.computeCurrent = false;
  cfg.computeFreeEnergy = false;
  std::unique_ptr<Setup> s = buildSetup(cfg);
  s->keldysh = true;
  s->spec = sp;
  setKeldyshBlock(*s, 0);
  return s;
}

std::unique_ptr<Setup> buildSetup(const Config &cfg) {
  auto sp = std::make_unique<Setup>();
  Setup &s = *sp;
  s.cfg = cfg;
  latticeSizes(cfg.grainWidth, cfg.ppxi, s.N0, s.N1, s.grainFraction);
  s.h = 1.0 / cfg.ppxi;
  s.temperature = cfg.temperature;
  s.lambda = cfg.lambda;
  s.Na = cfg.numAngles;
  s.Nc = cfg.numComponents;
  const int N0 = s.N0, N1 = s.N1, Na = s.Na;
  const double T = cfg.temperature;

  // energies
  s.Ne = numOzakiEnergies(T, cfg.cutoff);
  std::vector<double> poles;
  ozakiPoles(s.Ne, poles, s.residue);
  s.energy.resize(s.Ne);
  for (int i = 0; i < s.Ne; ++i) s.energy[i] = poles[i] * T / (2.0 * M_PI);
  double inv = 0.0;
  for (int i = 0; i < s.Ne; ++i) inv += s.residue[i] / s.energy[i];
  s.sumInverseEnergy = inv * T;

  // Fermi surface and rotation trigonometry
  s.polar.resize(Na);
  s.systemAngle.resize(Na);
  s.weight.resize(Na);
  s.vFx.resize(Na);
  s.vFy.resize(Na);
  s.saCos.resize(Na);
  s.saSin.resize(Na);
  s.opCos.resize(Na);
  s.opSin.resize(Na);
  s.aCos.resize(Na);
  s.aSin.resize(Na);
  s.bkCos.resize(Na);
  s.bkSin.resize(Na);
  for (int k = 0; k < Na; ++k) {
    s.polar[k] = polarAngle(k, Na);
    s.systemAngle[k] = s.polar[k] - 0.5 * M_PI;
    s.weight[k] = angleWeight(k, Na) / static_cast<double>(Na);
    s.vFx[k] = std::cos(s.polar[k]);
    s.vFy[k] = std::sin(s.polar[k]);
    s.saCos[k] = std::cos(s.systemAngle[k]);
    s.saSin[k] = std::sin(s.systemAngle[k]);
  }
  for (int k = 0; k < Na; ++k) {
    const float pOp = static_cast<float>(s.systemAngle[wrapIndex(-k, 0, Na)]);
    s.opCos[k] = std::cos(-pOp);
    s.opSin[k] = std::sin(-pOp);
    const float pA = static_cast<float>(-s.systemAngle[k]);
    s.aCos[k] = std::cos(-pA);
    s.aSin[k] = std::sin(-pA);
    const float pB = static_cast<float>(s.systemAngle[k]);
    s.bkCos[k] = std::cos(-pB);
    s.bkSin[k] = std::sin(-pB);
  }

  // order parameter basis, coupling constants, bulk values
  const double sBulk = bulkGapSWave(T);
  const double dOverS = std::exp(-kEulerGamma - 0.5) / std::sqrt(2.0) /
                        (0.5 * std::exp(-kEulerGamma));
  s.basis.resize(size_t(s.Nc) * Na);
  for (int c = 0; c < s.Nc; ++c) {
    double avg = 0.0;
    for (int k = 0; k < Na; ++k) {
      const double b = basisValue(cfg.symmetry[c], s.vFx[k], s.vFy[k]);
      avg += b * b * s.weight[k];
    }
    const double nrm = 1.0 / std::sqrt(avg);
    s.basisNorm.push_back(nrm);
    for (int k = 0; k < Na; ++k) {
      double rx, ry;
      rot(cfg.crystalRotation, s.vFx[k], s.vFy[k], rx, ry);
      s.basis[size_t(c) * Na + k] = basisValue(cfg.symmetry[c], rx, ry) * nrm;
    }
    s.coupling.push_back(0.5 * T / (std::log(T / cfg.tc[c]) + s.sumInverseEnergy));
    s.deltaBulk.push_back(cfg.symmetry[c] == SYM_S ? sBulk : sBulk * dOverS);
  }

  // reference-frame domain
  s.domain = referenceDomain(cfg.shapes, N0);
  s.numGrain = 0;
  for (size_t i = 0; i < s.domain.size(); ++i)
    if (s.domain[i]) {
      ++s.numGrain;
      s.domainSites.push_back(static_cast<int>(i));
    }
  if (s.numGrain == 0) throw std::runtime_error("the grain has no lattice sites");
  {
    std::vector<int8_t> lab(s.domain.begin(), s.domain.end());
    extensionList(lab.data(), N0, 0, s.extRef);
    s.extRef.offset.insert(s.extRef.offset.begin(), 0);
  }

  // rotational frames
  const size_t n1 = size_t(N1) * N1;
  s.labels.assign(size_t(Na) * n1, 0);
  struct Crossing {
    int y;
    double dist, nx, ny;
  };
  // per angle, per column: entries/exits along +y and -y
  struct FrameCross {
    std::vector<std::vector<Crossing>> inG, outG, inGt, outGt; // [x]
  };
  std::vector<FrameCross> fc(Na);
  s.colFirst.assign(size_t(Na) * N1, -1);
  s.colLast.assign(size_t(Na) * N1, -1);
  parallelFor(Na, [&](int k) {
    Frame f;
    groupFrame(cfg.shapes, -s.systemAngle[k], N0, N1, f);
    std::memcpy(&s.labels[size_t(k) * n1], f.label.data(), n1);
    FrameCross &x = fc[k];
    x.inG.resize(N1);
    x.outG.resize(N1);
    x.inGt.resize(N1);
    x.outGt.resize(N1);
    for (int col = 0; col < N1; ++col) {
      for (int y = 0; y < N1; ++y) {
        const size_t i = size_t(y) * N1 + col;
        const int8_t l = f.label[i];
        if (l > 0) {
          if (s.colFirst[size_t(k) * N1 + col] < 0) s.colFirst[size_t(k) * N1 + col] = y;
          s.colLast[size_t(k) * N1 + col] = y;
        }
        const Crossing cr{y, f.dist[i], f.nx[i], f.ny[i]};
        if (l == 2) {
          x.inG[col].push_back(cr);
          x.outGt[col].push_back(cr);
        } else if (l == 3) {
          x.outG[col].push_back(cr);
          x.inGt[col].push_back(cr);
        }
      }
      std::reverse(x.inGt[col].begin(), x.inGt[col].end());
      std::reverse(x.outGt[col].begin(), x.outGt[col].end());
    }
  });

  // sparse storage bookkeeping in (angle, column, trajectory order)
  s.idxInG.assign(size_t(Na) * N1, -1);
  s.idxOutG = s.idxInG;
  s.idxInGt = s.idxInG;
  s.idxOutGt = s.idxInG;
  s.numIsG.assign(size_t(Na) * N1, 0);
  s.numIsGt = s.numIsG;
  long long cIn = 0, cOut = 0, cInT = 0, cOutT = 0;
  for (int k = 0; k < Na; ++k)
    for (int col = 0; col < N1; ++col) {
      const size_t q = size_t(k) * N1 + col;
      const FrameCross &x = fc[k];
      if (!x.inG[col].empty()) s.idxInG[q] = static_cast<int>(cIn);
      cIn += x.inG[col].size();
      if (!x.outG[col].empty()) s.idxOutG[q] = static_cast<int>(cOut);
      cOut += x.outG[col].size();
      if (!x.inGt[col].empty()) s.idxInGt[q] = static_cast<int>(cInT);
      cInT += x.inGt[col].size();
      if (!x.outGt[col].empty()) s.idxOutGt[q] = static_cast<int>(cOutT);
      cOutT += x.outGt[col].size();
      s.numIsG[q] = static_cast<int>(x.outG[col].size());
      s.numIsGt[q] = static_cast<int>(x.outGt[col].size());
      for (const Crossing &cr : x.inG[col]) s.distInG.push_back(std::fabs(cr.dist));
      for (const Crossing &cr : x.outG[col]) {
        s.distOutG.push_back(std::fabs(cr.dist));
        s.coordG.push_back(cr.y);
        s.colOutG.push_back(static_cast<int>(q));
      }
      for (const Crossing &cr : x.inGt[col]) s.distInGt.push_back(std::fabs(cr.dist));
      for (const Crossing &cr : x.outGt[col]) {
        s.distOutGt.push_back(std::fabs(cr.dist));
        s.coordGt.push_back(cr.y);
        s.colOutGt.push_back(static_cast<int>(q));
      }
    }
  s.nInG = cIn;
  s.nOutG = cOut;
  s.nInGt = cInT;
  s.nOutGt = cOutT;
  if (cIn > std::numeric_limits<int>::max() / std::max(1, s.Ne))
    throw std::runtime_error("boundary storage too large");

  // trajectory ranges (BoundaryStorage.h:298-401)
  s.xmin.assign(Na, 0);
  s.xmax.assign(Na, 0);
  s.ymin.assign(Na, N1 - 1);
  s.ymax.assign(Na, 0);
  for (int k = 0; k < Na; ++k) {
    int a = -1, b = -1;
    for (int col = 0; col < N1; ++col)
      if (s.idxInG[size_t(k) * N1 + col] > -1) {
        if (a < 0) a = col;
        b = col;
      }
    s.xmin[k] = a < 0 ? 0 : a;
    s.xmax[k] = a < 0 ? 0 : b;
    for (int col = s.xmin[k]; col <= s.xmax[k]; ++col) {
      const int f = s.colFirst[size_t(k) * N1 + col];
      if (f >= 0) {
        s.ymin[k] = std::min(s.ymin[k], f);
        s.ymax[k] = std::max(s.ymax[k], s.colLast[size_t(k) * N1 + col]);
      }
    }
  }

  // specular-reflection stencils (BoundaryConditionSpecular.h:51-358)
  const double coordOff = static_cast<double>(N1 - 1) * 0.5;
  auto stencil = [&](bool gamma, int k, int x, int y, double bdist, double rmx,
                     double rmy, std::vector<int> &src, std::vector<double> &w,
                     double &wsum) {
    const int *idxOut = gamma ? s.idxOutG.data() : s.idxOutGt.data();
    const int *numIs = gamma ? s.numIsG.data() : s.numIsGt.data();
    const int *coord = gamma ? s.coordG.data() : s.coordGt.data();
    const double sa = s.systemAngle[k];
    double mx, my;
    rot(sa, rmx, rmy, mx, my);
    const double fai = fractionalAngleIndex(mx, my, Na);
    const int ai[2] = {wrapIndex(static_cast<int>(fai), 0, Na),
                       wrapIndex(static_cast<int>(fai) + 1, 0, Na)};
    const double frac = fai - std::floor(fai);
    const double aw[2] = {1.0 - frac, frac};
    const double px = static_cast<double>(x) - coordOff;
    const double py = static_cast<double>(y) -
                      static_cast<double>(gamma ? 1 : -1) * bdist - coordOff;
    src.clear();
    w.clear();
    wsum = 0.0;
    for (int r = 0; r < 2; ++r) {
      double qx, qy;
      rot(sa - s.systemAngle[ai[r]], px, py, qx, qy);
      qx += coordOff;
      qy += coordOff;
      const int pi[2] = {static_cast<int>(qx), static_cast<int>(qx) + 1};
      const double pf = qx - std::floor(qx);
      const double pw[2] = {1.0 - pf, pf};
      for (int t = 0; t < 2; ++t) {
        if (pi[t] < 0 || pi[t] >= N1) continue;
        const size_t q = size_t(ai[r]) * N1 + pi[t];
        int best = 0;
        double minDy = 1e6;
        for (int i = 0; i < numIs[q]; ++i) {
          const double dy = std::fabs(static_cast<double>(coord[idxOut[q] + i]) - qy);
          if (dy < minDy) {
            best = i;
            minDy = dy;
          }
        }
        if (idxOut[q] < 0) continue; // negative sparse index: no crossing
        const double ww = pw[t] * aw[r];
        src.push_back(idxOut[q] + best);
        w.push_back(ww);
        wsum += ww;
      }
    }
  };
  auto buildBc = [&](bool gamma, BcTable &tab) {
    const long long nIn = gamma ? s.nInG : s.nInGt;
    std::vector<std::vector<int>> srcs(nIn);
    std::vector<std::vector<double>> ws(nIn);
    tab.wsum.assign(nIn, 0.0);
    parallelFor(Na, [&](int k) {
      const double sa = s.systemAngle[k];
      double mx, my;
      rot(-sa, s.vFx[k], s.vFy[k], mx, my);
      const FrameCross &x = fc[k];
      for (int col = 0; col < N1; ++col) {
        const std::vector<Crossing> &ins = gamma ? x.inG[col] : x.inGt[col];
        if (ins.empty()) continue;
        const int base = (gamma ? s.idxInG : s.idxInGt)[size_t(k) * N1 + col];
        for (size_t e = 0; e < ins.size(); ++e) {
          const Crossing &cr = ins[e];
          const double bd = std::fabs(cr.dist);
          const double two = 2.0 * (mx * cr.nx + my * cr.ny);
          std::vector<int> src;
          std::vector<double> w;
          double wsum;
          stencil(gamma, k, col, cr.y, bd, mx - two * cr.nx, my - two * cr.ny, src,
                  w, wsum);
          if (wsum < 1e-7) stencil(gamma, k, col, cr.y, bd, -mx, -my, src, w, wsum);
          srcs[base + e] = src;
          ws[base + e] = w;
          tab.wsum[base + e] = wsum;
        }
      }
    });
    tab.offset.assign(nIn + 1, 0);
    for (long long i = 0; i < nIn; ++i) {
      tab.offset[i + 1] = tab.offset[i] + static_cast<int>(srcs[i].size());
      tab.src.insert(tab.src.end(), srcs[i].begin(), srcs[i].end());
      tab.w.insert(tab.w.end(), ws[i].begin(), ws[i].end());
    }
  };
  buildBc(true, s.bcG);
  buildBc(false, s.bcGt);

  // extension lists of the rotational frames (all angles)
  {
    std::vector<ExtensionList> per(Na);
    parallelFor(Na, [&](int k) {
      extensionList(&s.labels[size_t(k) * n1], N1, 0, per[k]);
    });
    s.extRot.offset.push_back(0);
    for (int k = 0; k < Na; ++k) {
      const long long shift = static_cast<long long>(k) * static_cast<long long>(n1);
      const int base = static_cast<int>(s.extRot.source.size());
      for (size_t t = 0; t < per[k].target.size(); ++t) {
        s.extRot.target.push_back(static_cast<int>(per[k].target[t] + shift));
        s.extRot.offset.push_back(base + per[k].offset[t]);
      }
      for (int v : per[k].source) s.extRot.source.push_back(static_cast<int>(v + shift));
    }
  }

  // site-field tiles: per frame, groups of 16 columns from xmin with the
  // union of the columns' [first - 1, last + 1]
  s.fieldTileStart.assign(size_t(Na) + 1, 0);
  for (int k = 0; k < Na; ++k) {
    s.fieldTileStart[size_t(k)] = static_cast<int>(s.fieldTiles.size() / 4);
    for (int x0 = s.xmin[k]; x0 <= s.xmax[k]; x0 += 16) {
      int lo = N1, hi = -1;
      for (int c = 0; c < 16 && x0 + c < N1; ++c) {
        const size_t q = size_t(k) * N1 + x0 + c;
        if (s.colFirst[q] < 0) continue;
        lo = std::min(lo, s.colFirst[q]);
        hi = std::max(hi, s.colLast[q]);
      }
      if (hi < 0) continue;
      s.fieldTiles.insert(s.fieldTiles.end(),
                          {k, x0, std::max(lo - 1, 0), std::min(hi + 1, N1 - 1)});
    }
  }
  s.fieldTileStart[size_t(Na)] = static_cast<int>(s.fieldTiles.size() / 4);

  long long inside = 0;
  for (int8_t l : s.labels) inside += l > 0;
  s.insideSites = inside;
  s.trajSteps = 2LL * s.Ne * inside;
  return sp;
}

std::vector<SweepItem> sweepItems(const Setup &s, int cols, int *maxLen) {
  std::vector<SweepItem> items;
  const int N1 = s.N1;
  int mx = 1;
  for (int k = 0; k < s.Na; ++k) {
    for (int x0 = s.xmin[k]; x0 <= s.xmax[k]; x0 += cols) {
      const int nc = std::min(cols, s.xmax[k] - x0 + 1);
      int lo = N1, hi = -1;
      for (int c = 0; c < nc; ++c) {
        const size_t q = size_t(k) * N1 + x0 + c;
        if (s.colFirst[q] < 0) continue;
        lo = std::min(lo, s.colFirst[q]);
        hi = std::max(hi, s.colLast[q]);
      }
      if (hi < 0) continue;
      // one extra site at each end so that the previous-site values of the
      // first inside site are available (the reference starts at ymin/ymax).
      lo = std::max(lo - 1, s.ymin[k]);
      hi = std::min(hi + 1, s.ymax[k]);
      items.push_back({k, x0, nc, lo, hi});
      mx = std::max(mx, hi - lo + 1);
    }
  }
  std::stable_sort(items.begin(), items.end(), [](const SweepItem &a, const SweepItem &b) {
    return (a.yhi - a.ylo) > (b.yhi - b.ylo);
  });
  if (maxLen) *maxLen = mx;
  return items;
}

} // namespace scamd
