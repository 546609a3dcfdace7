// sc_view.hpp -- the per-grain view the kernels read and the sweep's work
// description (no kernels: safe to include from host-side translation units).
#pragma once

#include <cstdint>

namespace scamd {

// Per-grain constants and device pointers consumed by the kernels.
struct GrainView {
  int N0, N1, Ne, Na, Nc;
  double h;
  double lambdaInv2;  // 1 / lambda^2, 0 when the induced potential is off
  double fluxDensity; // pi * Phi / area
  double area;
  int gauge;          // 0 landau, 1 symmetric, 2 solenoid
  int charge;
  int useInduced;     // lambda > 0
  int useField;       // any vector potential (induced or external)
  const double *energy, *residue; // Im eps_e (Matsubara / broadening), weight R_e
  const double *energyRe;         // Re eps_e for real-energy (Keldysh) sweeps, else null
  const float *opCos, *opSin, *aCos, *aSin, *bkCos, *bkSin;
  const double *saCos, *saSin, *vFx, *vFy, *weight, *basis;
  const int8_t *labels;
  const int *idxInG, *idxOutG, *idxInGt, *idxOutGt;
  const double *distInG, *distOutG, *distInGt, *distOutGt;
  double *bInG, *bOutG, *bInGt, *bOutGt; // planar [re: n*Ne][im: n*Ne]
  long long nInG, nOutG, nInGt, nOutGt;
  const double *extDelta; // [Nc][2][N0^2]
  const double *A;        // [2][N0^2] (scaled by lambda^2)
  double *G;              // [Na][N1^2][4]: Re/Im of f + f~*, Re g, Re i(D* g - D g~);
                          // Keldysh: [angle - angleBase][N1^2][Ne] spectral weight
  int angleBase;          // first angle held in G (Keldysh sweeps by angle chunks)
  // Site fields the sweep reads (Re D, Im D, Doppler shift), planar
  // [3][frames][N1^2] from frame fieldBase on, fieldPlane = frames * N1^2.
  const double *fields;
  long long fieldPlane;
  int fieldBase;
  // the tiles of swept sites the site fields cover, 4 ints each (frame,
  // first column, first row, last row), all frames
  const int *fieldTiles;
  int numFieldTiles;
  int cols;   // sweep lane layout: columns per item (64 / groups)
  int groups; // energy groups per column (ceil(Ne / EL)), one lane each
};

struct SweepItemDev {
  int grain, angle, x0, ncols, ylo, yhi;
};

// One launch sweeps the items of one or several grains (a batch); each item
// names its grain's view.
struct SweepArgs {
  const GrainView *views;
  const SweepItemDev *items;
  int numItems;
  int *counter;
  void *scratch;       // per slot: [maxLen][EL][64] complex T
  long long slotStride; // elements of complex T per slot
  int burnin;
};

} // namespace scamd
