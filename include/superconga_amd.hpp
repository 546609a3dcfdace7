/* superconga_amd.hpp -- C++ adaptors with the method sets of SuperConga's
 * solver interfaces, over the C ABI of superconga_amd.h.
 *
 * A maintainer who keeps the reference's C++ host code swaps its modules for
 * these classes: each one names the reference interface it stands in for
 * (file:line in the reference tree) and keeps its method names, argument
 * meaning and return values.  T is the reference's floating-point template
 * parameter: double runs the FP64 sweep, float the sweep with FP32 stepping
 * and FP64 sums; every field crosses the boundary in FP64.
 *
 * Ownership follows the reference's Context (Context.h:198-249): the
 * Context adaptor owns the device state (one sc_context); the other
 * adaptors borrow it and must not outlive it.
 *
 * Errors: the reference prints a failed device call and continues
 * (ContextModule.h:77-83) and exits on configuration errors; here a failed
 * call throws superconga_amd::Error carrying the library's message, which a
 * caller may print and continue from, or let end the program.
 */
#pragma once

#include "superconga_amd.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

namespace superconga_amd {

class Error : public std::runtime_error {
public:
  Error(int status, const std::string &what)
      : std::runtime_error(what), status_(status) {}
  int status() const { return status_; }

private:
  int status_;
};

inline void check(int status, const char *call) {
  if (status != SC_OK)
    throw Error(status, std::string(call) + ": " + sc_last_error());
}

/* The sweep precision of the reference's template parameter. */
template <typename T> struct Precision;
template <> struct Precision<double> {
  static constexpr int value = SC_PRECISION_FP64;
};
template <> struct Precision<float> {
  static constexpr int value = SC_PRECISION_MIXED;
};

/* ComputeResidual.h:29-106 on host arrays: |old - new| / max(|old|, eps) in
 * the chosen norm, eps = 1e-4 (times the element count for L1 and L2). */
enum class ResidualNorm { L1, L2, LInf };
inline double computeResidual(const double *oldValues, const double *newValues, size_t n,
                              ResidualNorm norm) {
  double numer = 0.0, denom = 0.0, eps = 1e-4;
  for (size_t i = 0; i < n; ++i) {
    const double d = std::fabs(oldValues[i] - newValues[i]), o = std::fabs(oldValues[i]);
    switch (norm) {
    case ResidualNorm::L1:
      numer += d;
      denom += o;
      break;
    case ResidualNorm::L2:
      numer += d * d;
      denom += o * o;
      break;
    case ResidualNorm::LInf:
      numer = std::max(numer, d);
      denom = std::max(denom, o);
      break;
    }
  }
  if (norm == ResidualNorm::L2) {
    numer = std::sqrt(numer);
    denom = std::sqrt(denom);
  }
  if (norm != ResidualNorm::LInf) eps *= static_cast<double>(n);
  return numer / std::max(denom, eps);
}

/* Context<T> (Context.h:25-511): one grain with its parameters, geometry,
 * order parameter, Riccati solver, compute objects and accelerator, built
 * from the text of a simulation_config.json (src/simulation.cu:64-463). */
template <typename T> class Context {
public:
  explicit Context(const std::string &configJson) {
    check(sc_create(configJson.c_str(), &ctx_), "sc_create");
    check(sc_get_dims(ctx_, &dims_), "sc_get_dims");
  }
  ~Context() {
    if (ctx_) sc_destroy(ctx_);
  }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;

  sc_context *handle() const { return ctx_; }
  const sc_dims &dims() const { return dims_; }
  int getGridResolutionBase() const { return dims_.n0; }
  int getNumComponents() const { return dims_.num_components; }
  int getNumFermiMomenta() const { return dims_.num_angles; }

  /* Context::compute (Context.h:391-447) and Context::burnIn (:377-389). */
  void compute() { check(sc_compute(ctx_, Precision<T>::value), "sc_compute"); }
  void burnIn() { check(sc_burnin(ctx_, Precision<T>::value), "sc_burnin"); }
  /* runSimulation's loop (run_simulation.h:124-187, run_iteration.h:85-141)
   * with the config's limits where an argument is negative. */
  bool run(int numBurnin = -1, int numMin = -1, int numMax = -1) {
    int converged = 0;
    check(sc_run(ctx_, numBurnin, numMin, numMax, Precision<T>::value, &converged), "sc_run");
    return converged != 0;
  }

  sc_results results() const {
    sc_results r;
    check(sc_get_results(ctx_, &r), "sc_get_results");
    return r;
  }
  /* Context::getResidual / isConverged / getNumIterations (Context.h:449-511). */
  T getResidual() const { return static_cast<T>(results().residual); }
  bool isConverged() const { return results().converged != 0; }
  int getNumIterations() const { return results().iterations; }
  T getFreeEnergyPerUnitArea() const { return static_cast<T>(results().free_energy_per_area); }

  /* OrderParameter::getComponentGrid (OrderParameter.h) of the iterate
   * (which = 0) or of the last gap equation's result (which = 1), and
   * OrderParameter::initialize. */
  std::vector<std::complex<double>> getOrderParameter(int component, int which = 0) const {
    const size_t n = size_t(dims_.n0) * dims_.n0;
    std::vector<double> re(n), im(n);
    check(sc_get_order_parameter(ctx_, which, component, re.data(), im.data()),
          "sc_get_order_parameter");
    std::vector<std::complex<double>> out(n);
    for (size_t i = 0; i < n; ++i) out[i] = {re[i], im[i]};
    return out;
  }
  void setOrderParameter(int component, const std::vector<std::complex<double>> &values) {
    const size_t n = size_t(dims_.n0) * dims_.n0;
    if (values.size() != n) throw Error(SC_ERR_ARGUMENT, "order parameter of the wrong size");
    std::vector<double> re(n), im(n);
    for (size_t i = 0; i < n; ++i) {
      re[i] = values[i].real();
      im[i] = values[i].imag();
    }
    check(sc_set_order_parameter(ctx_, component, re.data(), im.data()),
          "sc_set_order_parameter");
  }
  /* The induced vector potential, unscaled (the reference stores it times
   * lambda^2, RiccatiSolver.h:157-162): which = 0 iterate, 1 new solve. */
  void getVectorPotential(std::vector<double> &ax, std::vector<double> &ay,
                          int which = 0) const {
    const size_t n = size_t(dims_.n0) * dims_.n0;
    ax.resize(n);
    ay.resize(n);
    check(sc_get_vector_potential(ctx_, which, ax.data(), ay.data()),
          "sc_get_vector_potential");
  }
  void setVectorPotential(const std::vector<double> &ax, const std::vector<double> &ay) {
    const size_t n = size_t(dims_.n0) * dims_.n0;
    if (ax.size() != n || ay.size() != n)
      throw Error(SC_ERR_ARGUMENT, "vector potential of the wrong size");
    check(sc_set_vector_potential(ctx_, ax.data(), ay.data()), "sc_set_vector_potential");
  }
  std::vector<char> getDomain() const {
    std::vector<char> d(size_t(dims_.n0) * dims_.n0);
    check(sc_get_domain(ctx_, reinterpret_cast<signed char *>(d.data())), "sc_get_domain");
    return d;
  }

private:
  sc_context *ctx_ = nullptr;
  sc_dims dims_{};
};

/* BoundaryCondition<T> (boundary/BoundaryCondition.h:19-25) with the
 * specular reflection of BoundaryConditionSpecular<T>
 * (BoundaryConditionSpecular.h:361-486): out-values to in-values for every
 * angle and energy; returns the residual against the previous in-values. */
template <typename T> class BoundaryConditionSpecular {
public:
  T computeBoundaryCondition(Context<T> &ctx) {
    double r = 0.0;
    check(sc_boundary_condition(ctx.handle(), &r), "sc_boundary_condition");
    return static_cast<T>(r);
  }
};

/* RiccatiSolver<T> (riccati/RiccatiSolver.h:38-123) of a confined grain
 * (RiccatiSolverConfined<T, Gauge>, RiccatiSolverConfined.h:26-59): the
 * coherence functions of one angle for every Ozaki energy, between
 * sc_integration_begin and sc_integration_end (IntegrationIteratorOzaki
 * below drives it).  The Green-function sums of computeGreensAngle are
 * fused into the same sweep. */
template <typename T> class RiccatiSolver {
public:
  explicit RiccatiSolver(Context<T> &ctx) : ctx_(ctx) {}
  int initialize() { return 0; } /* tables are built with the context */
  void computeCoherenceFunctions(int angleIdx) {
    check(sc_sweep_angles(ctx_.handle(), Precision<T>::value, 0, angleIdx, angleIdx + 1),
          "sc_sweep_angles");
  }
  void computeCoherenceFunctionsBurnIn(int angleIdx) {
    check(sc_sweep_angles(ctx_.handle(), Precision<T>::value, 1, angleIdx, angleIdx + 1),
          "sc_sweep_angles");
  }
  /* RiccatiSolver::computeBoundaryConditions: the boundary condition it owns. */
  T computeBoundaryConditions() {
    residual_ = bc_.computeBoundaryCondition(ctx_);
    return residual_;
  }
  T getResidual() const { return residual_; }
  /* BoundaryStorage (BoundaryStorage.h:158-169): the sparse in-values
   * (inOut = 0) or out-values (1) of gamma (dir = 0) or gamma~ (1), laid
   * out (crossing * Ne + energy). */
  std::vector<std::complex<double>> getBoundaryStorage(int inOut, int dir) const {
    const sc_dims &d = ctx_.dims();
    const long long crossings = inOut == 0 ? (dir == 0 ? d.num_in_gamma : d.num_in_gammatilde)
                                           : (dir == 0 ? d.num_out_gamma : d.num_out_gammatilde);
    const size_t n = size_t(crossings) * size_t(d.num_energies);
    std::vector<double> re(n), im(n);
    check(sc_get_boundary(ctx_.handle(), inOut, dir, re.data(), im.data()), "sc_get_boundary");
    std::vector<std::complex<double>> out(n);
    for (size_t i = 0; i < n; ++i) out[i] = {re[i], im[i]};
    return out;
  }

private:
  Context<T> &ctx_;
  BoundaryConditionSpecular<T> bc_;
  T residual_ = 0;
};

/* ComputeProperty<T> (compute/ComputeProperty.h:27-103).  On the device the
 * order parameter, the current and the free energy are integrated by one
 * back-rotation per angle, which IntegrationIteratorOzaki issues once per
 * angle for every registered property; the hooks keep the reference's call
 * order and read the results back in computePropertyFromGreens. */
template <typename T> class ComputeProperty {
public:
  explicit ComputeProperty(Context<T> &ctx) : ctx_(ctx) {}
  virtual ~ComputeProperty() = default;
  virtual void initializeIteration() { angles_ = 0; }
  virtual void initializeGreensAngle() {}
  virtual void computeGreensAngle(int /*angleIdx*/) {} /* fused into the sweep */
  virtual void integrateGreensAngle(int /*angleIdx*/) { ++angles_; }
  virtual void computePropertyFromGreens() = 0;
  virtual T getResidual() const = 0;
  bool isConverged(T criterion) const { return getResidual() < criterion; }
  int integratedAngles() const { return angles_; }

protected:
  Context<T> &ctx_;
  int angles_ = 0;
};

/* ComputeOrderParameter<T> (compute/ComputeOrderParameter.h:174-262). */
template <typename T> class ComputeOrderParameter : public ComputeProperty<T> {
public:
  using ComputeProperty<T>::ComputeProperty;
  void computePropertyFromGreens() override {
    result_.clear();
    for (int c = 0; c < this->ctx_.getNumComponents(); ++c)
      result_.push_back(this->ctx_.getOrderParameter(c, 1));
  }
  const std::vector<std::vector<std::complex<double>>> &getResult() const { return result_; }
  T getResidual() const override {
    return static_cast<T>(this->ctx_.results().residual_order_parameter);
  }

private:
  std::vector<std::vector<std::complex<double>>> result_;
};

/* ComputeCurrent<T> (compute/ComputeCurrent.h:142-222). */
template <typename T> class ComputeCurrent : public ComputeProperty<T> {
public:
  using ComputeProperty<T>::ComputeProperty;
  void computePropertyFromGreens() override {
    const size_t n = size_t(this->ctx_.dims().n0) * this->ctx_.dims().n0;
    jx_.resize(n);
    jy_.resize(n);
    check(sc_get_current(this->ctx_.handle(), jx_.data(), jy_.data()), "sc_get_current");
  }
  const std::vector<double> &getResultX() const { return jx_; }
  const std::vector<double> &getResultY() const { return jy_; }
  T getResidual() const override { return static_cast<T>(this->ctx_.results().residual_current); }

private:
  std::vector<double> jx_, jy_;
};

/* ComputeFreeEnergy<T> (compute/ComputeFreeEnergy.h:260-414). */
template <typename T> class ComputeFreeEnergy : public ComputeProperty<T> {
public:
  using ComputeProperty<T>::ComputeProperty;
  void computePropertyFromGreens() override {
    const size_t n = size_t(this->ctx_.dims().n0) * this->ctx_.dims().n0;
    density_.resize(n);
    check(sc_get_free_energy_density(this->ctx_.handle(), density_.data()),
          "sc_get_free_energy_density");
    const sc_results r = this->ctx_.results();
    freeEnergy_ = r.free_energy;
    perArea_ = r.free_energy_per_area;
  }
  const std::vector<double> &getResult() const { return density_; }
  T getFreeEnergy() const { return static_cast<T>(freeEnergy_); }
  T getFreeEnergyPerUnitArea() const { return static_cast<T>(perArea_); }
  T getResidual() const override {
    return static_cast<T>(this->ctx_.results().residual_free_energy);
  }

private:
  std::vector<double> density_;
  double freeEnergy_ = 0.0, perArea_ = 0.0;
};

/* IntegrationIteratorOzaki<T> (integration/IntegrationIteratorOzaki.h,
 * IntegrationIterator.h:23-112): the loop over Fermi-surface angles of one
 * iteration, calling the Riccati solver and the compute objects' hooks in
 * the reference's order (IntegrationIteratorOzaki.h:300-420).  Equal to
 * Context::compute / burnIn bit for bit; anglesPerCall > 1 sweeps and
 * integrates that many angles per device call. */
template <typename T> class IntegrationIteratorOzaki {
public:
  IntegrationIteratorOzaki(Context<T> &ctx, RiccatiSolver<T> &solver)
      : ctx_(ctx), solver_(solver) {}
  void add(ComputeProperty<T> *property) { properties_.push_back(property); }

  void compute(bool isBurnIn, int anglesPerCall = 1) {
    const int na = ctx_.getNumFermiMomenta();
    const int step = std::max(anglesPerCall, 1);
    check(sc_integration_begin(ctx_.handle(), isBurnIn ? 1 : 0), "sc_integration_begin");
    if (!isBurnIn)
      for (ComputeProperty<T> *p : properties_) p->initializeIteration();
    for (int k0 = 0; k0 < na; k0 += step) {
      const int k1 = std::min(k0 + step, na);
      check(sc_sweep_angles(ctx_.handle(), Precision<T>::value, isBurnIn ? 1 : 0, k0, k1),
            "sc_sweep_angles");
      if (isBurnIn) continue;
      check(sc_integrate_angles(ctx_.handle(), k0, k1), "sc_integrate_angles");
      for (int k = k0; k < k1; ++k)
        for (ComputeProperty<T> *p : properties_) {
          p->initializeGreensAngle();
          p->computeGreensAngle(k);
          p->integrateGreensAngle(k);
        }
    }
    check(sc_integration_end(ctx_.handle(), isBurnIn ? 1 : 0), "sc_integration_end");
    if (!isBurnIn)
      for (ComputeProperty<T> *p : properties_) p->computePropertyFromGreens();
  }
  /* IntegrationIterator::getEnergies / getResidues / getNumEnergiesInterval:
   * the Ozaki energies (imaginary parts) and residues. */
  std::vector<double> getEnergies() const { return energies(true); }
  std::vector<double> getResidues() const { return energies(false); }
  int getNumEnergiesInterval() const { return ctx_.dims().num_energies; }

private:
  std::vector<double> energies(bool wantEnergies) const {
    const int ne = ctx_.dims().num_energies;
    std::vector<double> e(static_cast<size_t>(ne)), r(static_cast<size_t>(ne));
    check(sc_get_energies(ctx_.handle(), e.data(), r.data()), "sc_get_energies");
    return wantEnergies ? e : r;
  }
  Context<T> &ctx_;
  RiccatiSolver<T> &solver_;
  std::vector<ComputeProperty<T> *> properties_;
};

/* Accelerator<T> (accelerators/Accelerator.h:25-71): the context's
 * accelerator (name, rank, step sizes of the config) on caller-supplied
 * fields.  x and g are the concatenated field of accelerators/utils.h:
 * 58-186 (Re, Im of every component, then lambda^2 Ax, lambda^2 Ay with a
 * finite penetration depth), numFields() * N0^2 values each; x receives the
 * accelerated iterate.  Successive calls continue one acceleration history. */
template <typename T> class Accelerator {
public:
  explicit Accelerator(Context<T> &ctx) : ctx_(ctx) {}
  int numFields() const {
    int n = 0;
    check(sc_num_fields(ctx_.handle(), &n), "sc_num_fields");
    return n;
  }
  void compute(std::vector<double> &x, const std::vector<double> &g) {
    const size_t n = size_t(numFields()) * ctx_.dims().n0 * ctx_.dims().n0;
    if (x.size() != n || g.size() != n)
      throw Error(SC_ERR_ARGUMENT, "accelerator fields of the wrong size");
    check(sc_accelerator_step(ctx_.handle(), x.data(), g.data()), "sc_accelerator_step");
  }

private:
  Context<T> &ctx_;
};

/* ComputeVectorPotential<T> (compute/ComputeVectorPotential.h:143-201):
 * lambda^2 A = -h^2 (G * (j - mean_grain j)) with the exact lattice kernel
 * (maxError 0) or its eigen-truncation (maxError > 0,
 * ComputeVectorPotential.h:319-395), on host arrays. */
template <typename T> class ComputeVectorPotential {
public:
  int initialize(int gridResolution, T latticeElementSize, int numGrainLatticeSites,
                 T maxError = 0) {
    n_ = gridResolution;
    h_ = latticeElementSize;
    numGrain_ = numGrainLatticeSites;
    maxError_ = maxError;
    return (n_ >= 2 && h_ > 0 && numGrain_ >= 1) ? 0 : -1;
  }
  /* solve(j): jx, jy of n^2 values each (the current's GridData planes). */
  void solve(const std::vector<double> &jx, const std::vector<double> &jy) {
    const size_t n = size_t(n_) * n_;
    if (jx.size() != n || jy.size() != n)
      throw Error(SC_ERR_ARGUMENT, "current of the wrong size");
    std::vector<double> j(2 * n);
    std::copy(jx.begin(), jx.end(), j.begin());
    std::copy(jy.begin(), jy.end(), j.begin() + n);
    result_.assign(2 * n, 0.0);
    check(sc_poisson_solve_ex(n_, double(h_), numGrain_, 1, double(maxError_), j.data(),
                              result_.data()),
          "sc_poisson_solve_ex");
  }
  /* The residual of the result against the current potential (both scaled
   * by lambda^2, 2 n^2 values). */
  T computeResidual(const std::vector<double> &current, ResidualNorm norm) {
    if (current.size() != result_.size())
      throw Error(SC_ERR_ARGUMENT, "vector potential of the wrong size");
    residual_ = static_cast<T>(
        superconga_amd::computeResidual(current.data(), result_.data(), current.size(), norm));
    return residual_;
  }
  const std::vector<double> &getResult() const { return result_; }
  T getResidual() const { return residual_; }

private:
  int n_ = 0, numGrain_ = 0;
  T h_ = 0, maxError_ = 0, residual_ = 0;
  std::vector<double> result_;
};

} // namespace superconga_amd
