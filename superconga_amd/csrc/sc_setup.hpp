// sc_setup.hpp -- iteration-invariant host data for one grain.
//
// Everything the reference recomputes on every angle of every iteration but
// which never changes (geometry labels, boundary storage bookkeeping,
// specular-reflection interpolation stencils, extension stencils, Ozaki poles,
// rotation trigonometry) is computed once here in FP64 on the host and
// uploaded.  The references to the reference's files state which routine each
// table replaces.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace scamd {

enum Symmetry { SYM_S = 0, SYM_DX2Y2 = 1, SYM_DXY = 2, SYM_G = 3 };
enum Gauge { GAUGE_LANDAU = 0, GAUGE_SYMMETRIC = 1, GAUGE_SOLENOID = 2 };
enum Norm { NORM_L1 = 0, NORM_L2 = 1, NORM_LINF = 2 };
enum Accel {
  ACC_PICARD = 0,
  ACC_ANDERSON = 1,
  ACC_CONGACC = 2,
  ACC_BB = 3,
  ACC_POLYAK = 4,
  // Broyden's second method, restarted limited memory (no counterpart in the
  // reference's accelerators/)
  ACC_BROYDEN = 5
};

constexpr int kMaxComponents = 4;

// A geometry component in the internal frame, already scaled by the grain
// fraction (DiscGeometry.h:249-261, PolygonGeometry.h:265-342).
struct Shape {
  bool disc = true;
  bool add = true;
  double radius = 0.5, cx = 0.0, cy = 0.0; // disc
  std::vector<double> vx, vy;              // polygon vertices
};

struct Vortex {
  double cx, cy, winding;
};

struct Config {
  // physics (src/simulation.cu:101-155)
  double temperature = 0.5;
  double fluxQuanta = 0.0;
  double lambda = -1.0;
  double crystalRotation = 0.0; // radians
  int gauge = GAUGE_SYMMETRIC;
  int chargeSign = -1;
  int numComponents = 0;
  int symmetry[kMaxComponents] = {0, 0, 0, 0};
  double tc[kMaxComponents] = {1, 1, 1, 1};
  double phaseShift[kMaxComponents] = {0, 0, 0, 0};
  double noise[kMaxComponents] = {0, 0, 0, 0};
  std::vector<Vortex> vortices[kMaxComponents];
  // geometry
  double grainWidth = 20.0;
  double geomOffX = 0.0, geomOffY = 0.0; // bounding-box centre (user units)
  std::vector<Shape> shapes;
  // numerics
  double convergence = 1e-6;
  double cutoff = 10.0;
  int norm = NORM_L2;
  int numAngles = 64;
  double ppxi = 5.0;
  double vpError = 0.0;
  int burnin = 0, itMin = 0, itMax = 10000;
  bool computeCurrent = true, computeFreeEnergy = true;
  // accelerator (accelerators/*.h defaults)
  int accel = ACC_CONGACC;
  double stepSize = -1, drag = -1, stepFactor = -1, cosMin = -1, stepMax = -1,
         stepMin = -1;
  int rank = -1;
  // Anderson extension: weight each field block (a component of Delta, the
  // vector potential) by 1 / |x_block|^2 in the least squares
  bool blockScaling = false;
  // Extension: precondition the vector-potential update by the Meissner
  // response of a superfluid density of this value (0: off)
  double vpScreening = 0.0;
  // misc
  std::string loadPath;
  unsigned long long seed = 0;
  bool verbose = false;
};

// Parses simulation_config.json text.  Throws std::runtime_error.
Config parseConfig(const std::string &text);

// Compressed lists "target <- mean of sources" for extendValuesAtBoundary
// (geometry/domain_ops.h:11-57).  Sources are listed in the reference's scan
// order so that the sums are formed in the same order.
struct ExtensionList {
  std::vector<int> target; // flat site index
  std::vector<int> offset; // size target.size() + 1
  std::vector<int> source; // flat site index
};

// Specular-reflection stencil per boundary entry crossing
// (BoundaryConditionSpecular.h:108-358): in = sum_j w_j out[src_j] / wsum.
struct BcTable {
  std::vector<int> offset; // size numIn + 1
  std::vector<int> src;    // crossing index into the out storage
  std::vector<double> w;
  std::vector<double> wsum;
};

// One unit of work of the Riccati sweep: a block of adjacent columns of one
// rotational frame and the rows the block's trajectories touch.
struct SweepItem {
  int angle;
  int x0;
  int ncols;
  int ylo, yhi;
};

// postprocess_config.json (src/postprocess.cu:118-160,
// frontend/postprocess/arguments.py:309-345): real energies for the LDOS.
struct Spectroscopy {
  double eMin = 0.0, eMax = 0.0, broadening = 0.0;
  int numEnergies = 0;
  int perBlock = -1; // < 1: every energy in one block
  int numAngles = 0;
  double convergence = 1e-4;
  int norm = NORM_L2;
  int burnin = -1, itMin = 0, itMax = -1;
};
Spectroscopy parseSpectroscopy(const std::string &text);

struct Setup {
  Config cfg;
  // Keldysh spectroscopy: the energies are E + i broadening of one block
  // (energyRe, energy), the residues are one; spec holds the whole range.
  bool keldysh = false;
  Spectroscopy spec;
  std::vector<double> energyRe;
  long long insideSites = 0; // #{sites with label > 0 over all frames}
  int N0 = 0, N1 = 0, Ne = 0, Na = 0, Nc = 0, numGrain = 0;
  double h = 0, grainFraction = 0, temperature = 0, lambda = -1;
  // energies (IntegrationIteratorOzaki.h:108-200)
  std::vector<double> energy, residue;
  // Fermi surface (FermiSurfaceCircle.h:83-163)
  std::vector<double> systemAngle, polar, weight, vFx, vFy, saCos, saSin;
  // single-precision rotation trigonometry of insertRotatedGrid(Additive)
  // (GridFunctions_GPU.h:436-544): cos(-phi), sin(-phi) with phi a float.
  std::vector<float> opCos, opSin; // order parameter into frame k
  std::vector<float> aCos, aSin;   // vector potential into frame k
  std::vector<float> bkCos, bkSin; // Green functions back from frame k
  // order parameter (OrderParameterBasis.h, OrderParameterComponent.h)
  std::vector<double> basis; // [c * Na + k]
  std::vector<double> basisNorm, coupling, deltaBulk;
  double sumInverseEnergy = 0; // T * sum_n R_n / Im eps_n
  // geometry
  std::vector<int8_t> labels;  // [Na][N1][N1]
  std::vector<uint8_t> domain; // [N0][N0] reference-frame labels (0/1)
  std::vector<int> domainSites;
  // boundary storage (BoundaryStorage.h:298-527)
  std::vector<int> idxInG, idxOutG, idxInGt, idxOutGt, numIsG, numIsGt;
  long long nInG = 0, nOutG = 0, nInGt = 0, nOutGt = 0;
  std::vector<int> coordG, coordGt;       // y of every out crossing (storage order)
  std::vector<int> colOutG, colOutGt;     // k * N1 + column of every out crossing
  std::vector<double> distInG, distOutG, distInGt, distOutGt;
  std::vector<int> xmin, xmax, ymin, ymax;
  std::vector<int> colFirst, colLast; // [Na][N1], -1 where empty
  // Site-field tiles: 16 adjacent columns of one frame and the rows a sweep
  // may read there (k, x0, y0, y1 per tile, frames in order);
  // fieldTileStart[k] is the first tile of frame k.
  std::vector<int> fieldTiles, fieldTileStart;
  BcTable bcG, bcGt;
  ExtensionList extRef, extRot;
  long long trajSteps = 0; // 2 * Ne * #{sites with label > 0 over all frames}
};

std::unique_ptr<Setup> buildSetup(const Config &cfg);
// The postprocess context (src/postprocess.cu:118-330): the simulation's
// physics and geometry with the spectroscopy's angles, criterion and norm,
// energies set to the first block.
std::unique_ptr<Setup> buildSpectroscopySetup(Config cfg, const Spectroscopy &sp);
// Energies [start, start + count) of the Keldysh grid into s (energy,
// energyRe, residue, Ne, trajSteps).
void setKeldyshBlock(Setup &s, int start);

// Work list of the sweep for a given number of columns per item, sorted by
// decreasing length.
std::vector<SweepItem> sweepItems(const Setup &s, int colsPerItem,
                                  int *maxLength);

// Stand-alone pieces exposed through the C ABI for unit checks.
int numOzakiEnergies(double temperature, double cutoff);
void ozakiPoles(int numPoles, std::vector<double> &poles,
                std::vector<double> &residues);
double bulkGapSWave(double temperature);
void latticeSizes(double width, double ppxi, int &n0, int &n1,
                  double &grainFraction);
// Real, even Green-function kernel of ComputeVectorPotential
// (ComputeVectorPotential.h:212-294,319-415): G(|dx|, |dy|) for
// 0 <= dx, dy < n as g[dy * n + dx].  maxError = 0 gives the exact kernel
// (returns 0); maxError > 0 the eigen-truncated one the reference's
// separable filters sum to (returns the number of rank-one filters,
// the constant included).
int poissonKernel(int n, double h, std::vector<double> &g, double maxError = 0.0);
// Eigenvalues and eigenvectors (rows of vecs) of a real symmetric n x n
// matrix given row-major in a (overwritten).
void symmetricEigen(int n, std::vector<double> &a, std::vector<double> &vals,
                    std::vector<double> &vecs);

} // namespace scamd
