// sc_capi.hip -- the extern "C" boundary of libsuperconga_amd.so.

#include "../../include/superconga_amd.h"
#include "sc_engine.hpp"
#include "sc_device.hpp"

#include <algorithm>
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

struct sc_context {
  std::unique_ptr<scamd::Grain> owned;
  scamd::Grain *grain = nullptr;
  bool borrowed = false; // a grain of a batch: the batch owns it
};

struct sc_batch {
  std::unique_ptr<scamd::Batch> batch;
  std::vector<std::unique_ptr<sc_context>> views;
};

namespace {
thread_local std::string g_lastError;

int fail(int code, const std::string &msg) {
  g_lastError = msg;
  return code;
}

template <class F> int guarded(F &&f) {
  try {
    return f();
  } catch (const std::invalid_argument &e) {
    return fail(SC_ERR_ARGUMENT, e.what());
  } catch (const std::runtime_error &e) {
    const std::string m = e.what();
    return fail(m.rfind("HIP error", 0) == 0 ? SC_ERR_DEVICE : SC_ERR_CONFIG, m);
  } catch (const std::exception &e) {
    return fail(SC_ERR_STATE, e.what());
  }
}

#define SC_CHECK_CTX(ctx)                                                      \
  if (!(ctx) || !(ctx)->grain) return fail(SC_ERR_ARGUMENT, "null context")
#define SC_CHECK_PTR(p)                                                        \
  if (!(p)) return fail(SC_ERR_ARGUMENT, "null pointer argument: " #p)
#define SC_CHECK_SIM(ctx)                                                      \
  if ((ctx)->grain->keldysh())                                                 \
  return fail(SC_ERR_STATE, "a spectroscopy context only runs sc_ldos_*")
#define SC_CHECK_LDOS(ctx)                                                     \
  if (!(ctx)->grain->keldysh())                                                \
  return fail(SC_ERR_STATE, "not a spectroscopy context (use sc_ldos_create)")


void checkPrecision(int p) {
  if (p != SC_PRECISION_FP64 && p != SC_PRECISION_MIXED)
    throw std::invalid_argument("precision must be SC_PRECISION_FP64 or SC_PRECISION_MIXED");
}

__global__ void riccatiStepsKernel(int n, const double *in, double *out, int prec) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *r = in + 7 * size_t(i);
  if (prec == SC_PRECISION_FP64) {
    const scamd::Cx<double> y = scamd::riccatiStep(scamd::Cx<double>{r[0], r[1]},
                                                   scamd::Cx<double>{r[2], r[3]},
                                                   scamd::Cx<double>{r[4], r[5]}, r[6]);
    out[2 * i] = y.re;
    out[2 * i + 1] = y.im;
  } else {
    const scamd::Cx<float> y = scamd::riccatiStep(
        scamd::Cx<float>{float(r[0]), float(r[1])}, scamd::Cx<float>{float(r[2]), float(r[3])},
        scamd::Cx<float>{float(r[4]), float(r[5])}, float(r[6]));
    out[2 * i] = y.re;
    out[2 * i + 1] = y.im;
  }
}

void hipCheck(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

} // namespace

namespace scamd {
int reportError(int code, const std::string &message) { return fail(code, message); }
} // namespace scamd

extern "C" {

const char *sc_last_error(void) { return g_lastError.c_str(); }
int sc_version(void) { return 1; }

int sc_create(const char *config_json, sc_context **out) {
  SC_CHECK_PTR(config_json);
  SC_CHECK_PTR(out);
  *out = nullptr;
  return guarded([&] {
    scamd::Config cfg = scamd::parseConfig(config_json);
    auto setup = scamd::buildSetup(cfg);
    auto ctx = std::make_unique<sc_context>();
    ctx->owned = std::make_unique<scamd::Grain>(std::move(setup));
    ctx->grain = ctx->owned.get();
    ctx->grain->initialGuess();
    *out = ctx.release();
    return SC_OK;
  });
}

int sc_ldos_create(const char *simulation_json, const char *postprocess_json, sc_context **out) {
  SC_CHECK_PTR(simulation_json);
  SC_CHECK_PTR(postprocess_json);
  SC_CHECK_PTR(out);
  *out = nullptr;
  return guarded([&] {
    scamd::Config cfg = scamd::parseConfig(simulation_json);
    const scamd::Spectroscopy sp = scamd::parseSpectroscopy(postprocess_json);
    auto setup = scamd::buildSpectroscopySetup(cfg, sp);
    auto ctx = std::make_unique<sc_context>();
    ctx->owned = std::make_unique<scamd::Grain>(std::move(setup));
    ctx->grain = ctx->owned.get();
    *out = ctx.release();
    return SC_OK;
  });
}

int sc_ldos_compute(sc_context *ctx, int precision, int burnin) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_LDOS(ctx);
  return guarded([&] {
    checkPrecision(precision);
    ctx->grain->ldosCompute(precision, burnin != 0);
    return SC_OK;
  });
}

int sc_ldos_run(sc_context *ctx, int nb, int nmin, int nmax, int precision, int *converged) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_LDOS(ctx);
  return guarded([&] {
    checkPrecision(precision);
    const bool ok = ctx->grain->ldosRun(nb, nmin, nmax, precision);
    if (converged) *converged = ok ? 1 : 0;
    return SC_OK;
  });
}

int sc_ldos_get_status(sc_context *ctx, sc_ldos_status *out) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_LDOS(ctx);
  SC_CHECK_PTR(out);
  return guarded([&] {
    const scamd::LdosStatus st = ctx->grain->ldosStatus();
    out->num_energies = st.numEnergies;
    out->energies_per_block = st.perBlock;
    out->block_start = st.blockStart;
    out->block_count = st.blockCount;
    out->iterations = st.iterations;
    out->angle_chunk = st.angleChunk;
    out->partially_converged = st.partiallyConverged ? 1 : 0;
    out->converged = st.converged ? 1 : 0;
    out->residual = st.residual;
    out->residual_boundary = st.residualBoundary;
    out->residual_ldos = st.residualLdos;
    out->last_sweep_ms = st.sweepMs;
    out->last_iteration_ms = st.iterationMs;
    out->traj_steps = st.trajSteps;
    return SC_OK;
  });
}

int sc_ldos_get(sc_context *ctx, double *out) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_LDOS(ctx);
  SC_CHECK_PTR(out);
  return guarded([&] {
    ctx->grain->getLdos(out);
    return SC_OK;
  });
}

int sc_ldos_set_angle_chunk(sc_context *ctx, int angles) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_LDOS(ctx);
  return guarded([&] {
    if (angles < 0) throw std::invalid_argument("angles must be >= 0");
    ctx->grain->setLdosAngleChunk(angles);
    return SC_OK;
  });
}

int sc_destroy(sc_context *ctx) {
  if (!ctx) return SC_OK;
  if (ctx->borrowed)
    return fail(SC_ERR_ARGUMENT, "a grain of a batch is destroyed with its batch");
  return guarded([&] {
    delete ctx;
    return SC_OK;
  });
}

int sc_get_dims(sc_context *ctx, sc_dims *out) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_PTR(out);
  const scamd::Setup &s = ctx->grain->setup();
  out->n0 = s.N0;
  out->n1 = s.N1;
  out->num_energies = s.Ne;
  out->num_angles = s.Na;
  out->num_components = s.Nc;
  out->num_grain_sites = s.numGrain;
  out->num_in_gamma = s.nInG;
  out->num_out_gamma = s.nOutG;
  out->num_in_gammatilde = s.nInGt;
  out->num_out_gammatilde = s.nOutGt;
  out->traj_steps = s.trajSteps;
  out->h = s.h;
  out->grain_fraction = s.grainFraction;
  return SC_OK;
}

int sc_get_energies(sc_context *ctx, double *e, double *r) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_PTR(e);
  SC_CHECK_PTR(r);
  const scamd::Setup &s = ctx->grain->setup();
  std::memcpy(e, s.energy.data(), s.energy.size() * sizeof(double));
  std::memcpy(r, s.residue.data(), s.residue.size() * sizeof(double));
  return SC_OK;
}

int sc_set_order_parameter(sc_context *ctx, int comp, const double *re, const double *im) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_PTR(re);
  SC_CHECK_PTR(im);
  return guarded([&] {
    ctx->grain->setOrderParameter(comp, re, im);
    return SC_OK;
  });
}
int sc_get_order_parameter(sc_context *ctx, int which, int comp, double *re, double *im) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_PTR(re);
  SC_CHECK_PTR(im);
  if (which != 0 && which != 1) return fail(SC_ERR_ARGUMENT, "which must be 0 or 1");
  return guarded([&] {
    ctx->grain->getOrderParameter(which, comp, re, im);
    return SC_OK;
  });
}
int sc_set_vector_potential(sc_context *ctx, const double *ax, const double *ay) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_PTR(ax);
  SC_CHECK_PTR(ay);
  return guarded([&] {
    ctx->grain->setVectorPotential(ax, ay, false);
    return SC_OK;
  });
}
int sc_get_vector_potential(sc_context *ctx, int which, double *ax, double *ay) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_PTR(ax);
  SC_CHECK_PTR(ay);
  if (which != 0 && which != 1) return fail(SC_ERR_ARGUMENT, "which must be 0 or 1");
  return guarded([&] {
    ctx->grain->getVectorPotential(which, ax, ay, false);
    return SC_OK;
  });
}
int sc_get_current(sc_context *ctx, double *jx, double *jy) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_PTR(jx);
  SC_CHECK_PTR(jy);
  return guarded([&] {
    ctx->grain->getCurrent(jx, jy);
    return SC_OK;
  });
}
int sc_get_free_energy_density(sc_context *ctx, double *out) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_PTR(out);
  return guarded([&] {
    ctx->grain->getEnergyDensity(out);
    return SC_OK;
  });
}
int sc_get_boundary(sc_context *ctx, int in_out, int dir, double *re, double *im) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_PTR(re);
  SC_CHECK_PTR(im);
  if ((in_out != 0 && in_out != 1) || (dir != 0 && dir != 1))
    return fail(SC_ERR_ARGUMENT, "in_out and dir must be 0 or 1");
  return guarded([&] {
    ctx->grain->getBoundary(in_out, dir, re, im);
    return SC_OK;
  });
}
int sc_set_boundary(sc_context *ctx, int dir, const double *re, const double *im) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_PTR(re);
  SC_CHECK_PTR(im);
  if (dir != 0 && dir != 1) return fail(SC_ERR_ARGUMENT, "dir must be 0 or 1");
  return guarded([&] {
    ctx->grain->setBoundary(dir, re, im);
    return SC_OK;
  });
}

int sc_compute(sc_context *ctx, int precision) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_SIM(ctx);
  return guarded([&] {
    checkPrecision(precision);
    ctx->grain->compute(precision);
    return SC_OK;
  });
}
int sc_burnin(sc_context *ctx, int precision) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_SIM(ctx);
  return guarded([&] {
    checkPrecision(precision);
    ctx->grain->burnin(precision);
    return SC_OK;
  });
}
int sc_run(sc_context *ctx, int nb, int nmin, int nmax, int precision, int *converged) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_SIM(ctx);
  return guarded([&] {
    if (precision != SC_PRECISION_MIXED_REFINE) checkPrecision(precision);
    const scamd::Config &c = ctx->grain->setup().cfg;
    const bool ok = ctx->grain->run(nb < 0 ? c.burnin : nb, nmin < 0 ? c.itMin : nmin,
                                    nmax < 0 ? c.itMax : nmax, precision);
    if (converged) *converged = ok ? 1 : 0;
    return SC_OK;
  });
}
int sc_get_results(sc_context *ctx, sc_results *out) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_PTR(out);
  const scamd::Results r = ctx->grain->results();
  const scamd::Timing t = ctx->grain->timing();
  out->residual = r.residual;
  out->mixed_iterations = ctx->grain->mixedIterations();
  out->residual_boundary = r.resBoundary;
  out->residual_order_parameter = r.resDelta;
  out->residual_current = r.resCurrent;
  out->residual_free_energy = r.resFreeEnergy;
  out->residual_vector_potential = r.resVectorPotential;
  out->free_energy = r.freeEnergy;
  out->free_energy_per_area = r.freeEnergyPerArea;
  out->magnetic_free_energy = r.magneticFreeEnergy;
  out->magnetic_free_energy_per_area = r.magneticFreeEnergyPerArea;
  out->iterations = r.iterations;
  out->converged = r.converged;
  out->last_sweep_ms = t.sweepMs;
  out->last_iteration_ms = t.iterMs;
  return SC_OK;
}

int sc_riccati_sweep(sc_context *ctx, int precision, int burnin) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_SIM(ctx);
  return guarded([&] {
    checkPrecision(precision);
    ctx->grain->sweep(precision, burnin != 0);
    hipCheck(hipStreamSynchronize(ctx->grain->stream()), "sweep");
    ctx->grain->updateSweepTime();
    return SC_OK;
  });
}
int sc_boundary_condition(sc_context *ctx, double *residual) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_SIM(ctx);
  return guarded([&] {
    ctx->grain->boundaryCondition();
    if (residual) *residual = ctx->grain->results().resBoundary;
    return SC_OK;
  });
}
int sc_set_sweep_shape(sc_context *ctx, int precision, int el, int waves, int slots) {
  SC_CHECK_CTX(ctx);
  return guarded([&] {
    checkPrecision(precision);
    ctx->grain->setSweepShape(precision, el, waves, slots);
    return SC_OK;
  });
}
int sc_set_screening(sc_context *ctx, double superfluidDensity) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_SIM(ctx);
  return guarded([&] {
    ctx->grain->setScreening(superfluidDensity);
    return SC_OK;
  });
}
int sc_set_graphs(sc_context *ctx, int enable) {
  SC_CHECK_CTX(ctx);
  return guarded([&] {
    ctx->grain->setGraphs(enable != 0);
    return SC_OK;
  });
}
int sc_get_sweep_shape(sc_context *ctx, int precision, int *el, int *waves, int *slots) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_PTR(el);
  SC_CHECK_PTR(waves);
  SC_CHECK_PTR(slots);
  return guarded([&] {
    checkPrecision(precision);
    ctx->grain->sweepShape(precision, *el, *waves, *slots);
    return SC_OK;
  });
}
int sc_autotune_sweep(sc_context *ctx, int precision) {
  SC_CHECK_CTX(ctx);
  return guarded([&] {
    checkPrecision(precision);
    ctx->grain->autotune(precision);
    hipCheck(hipStreamSynchronize(ctx->grain->stream()), "autotune");
    return SC_OK;
  });
}

int sc_newton_krylov(sc_context *ctx, int maxNewton, int krylovDim, double tol,
                     sc_newton_report *report) {
  return sc_newton_krylov_ex(ctx, maxNewton, krylovDim, tol, SC_PRECISION_FP64, report);
}

int sc_newton_krylov_ex(sc_context *ctx, int maxNewton, int krylovDim, double tol,
                        int jvpPrecision, sc_newton_report *report) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_SIM(ctx);
  return guarded([&] {
    if (maxNewton < 0) throw std::invalid_argument("max_newton must be >= 0");
    if (krylovDim < 1 || krylovDim > 63)
      throw std::invalid_argument("krylov_dim must be in [1, 63]");
    if (jvpPrecision != SC_PRECISION_FP64 && jvpPrecision != SC_PRECISION_MIXED &&
        jvpPrecision != SC_JVP_TANGENT)
      throw std::invalid_argument(
          "jvp_precision must be SC_PRECISION_FP64, SC_PRECISION_MIXED or SC_JVP_TANGENT");
    if (!(tol > 0.0)) throw std::invalid_argument("tol must be positive");
    if (report && report->history_capacity > 0 && !report->history)
      throw std::invalid_argument("history is null with a positive capacity");
    const scamd::NewtonReport r = ctx->grain->newtonKrylov(
        maxNewton, krylovDim, tol,
        jvpPrecision == SC_PRECISION_MIXED
            ? scamd::PREC_MIXED
            : (jvpPrecision == SC_JVP_TANGENT ? scamd::JVP_TANGENT : scamd::PREC_FP64));
    if (report) {
      report->newton_iterations = r.newtonIterations;
      report->krylov_iterations = r.krylovIterations;
      report->map_evaluations = r.mapEvaluations;
      report->converged = r.converged ? 1 : 0;
      report->residual = r.residual;
      const int n = std::min(static_cast<int>(r.history.size()),
                             std::max(report->history_capacity, 0));
      for (int i = 0; i < n; ++i) report->history[i] = r.history[static_cast<size_t>(i)];
      report->history_length = n;
    }
    return SC_OK;
  });
}

int sc_free_energy_gradient(sc_context *ctx, double *grad, double *freeEnergy) {
  return sc_free_energy_gradient_ex(ctx, grad, nullptr, freeEnergy);
}

int sc_set_newton_options(sc_context *ctx, int restarts, double etaMax, int fp64Fallback,
                          int fixPhase) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_SIM(ctx);
  return guarded([&] {
    ctx->grain->setNewtonOptions(restarts, etaMax, fp64Fallback != 0, fixPhase != 0);
    return SC_OK;
  });
}

int sc_set_adjoint_checkpoints(sc_context *ctx, int k) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_SIM(ctx);
  if (k < -1) return fail(SC_ERR_ARGUMENT, "k must be -1 (automatic), 0 (every site) or positive");
  return guarded([&] {
    ctx->grain->setAdjointCheckpoints(k);
    return SC_OK;
  });
}

int sc_jacobian_vector(sc_context *ctx, const double *v, double *jv) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_SIM(ctx);
  SC_CHECK_PTR(v);
  SC_CHECK_PTR(jv);
  return guarded([&] {
    ctx->grain->jacobianVectorHost(v, jv);
    return SC_OK;
  });
}

int sc_residual_gradient(sc_context *ctx, double *grad, double *gradA, double *residual) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_SIM(ctx);
  SC_CHECK_PTR(grad);
  return guarded([&] {
    const double r = ctx->grain->residualGradient(grad, gradA);
    if (residual) *residual = r;
    return SC_OK;
  });
}

int sc_free_energy_gradient_ex(sc_context *ctx, double *grad, double *gradA,
                               double *freeEnergy) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_SIM(ctx);
  SC_CHECK_PTR(grad);
  return guarded([&] {
    const double f = ctx->grain->freeEnergyGradient(grad, gradA);
    if (freeEnergy) *freeEnergy = f;
    return SC_OK;
  });
}

int sc_integration_begin(sc_context *ctx, int burnin) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_SIM(ctx);
  return guarded([&] {
    ctx->grain->beginIntegration(burnin != 0);
    return SC_OK;
  });
}

int sc_sweep_angles(sc_context *ctx, int precision, int burnin, int angleBegin, int angleEnd) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_SIM(ctx);
  return guarded([&] {
    checkPrecision(precision);
    ctx->grain->sweepAngles(precision, burnin != 0, angleBegin, angleEnd);
    return SC_OK;
  });
}

int sc_integrate_angles(sc_context *ctx, int angleBegin, int angleEnd) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_SIM(ctx);
  return guarded([&] {
    ctx->grain->integrateAngles(angleBegin, angleEnd);
    return SC_OK;
  });
}

int sc_integration_end(sc_context *ctx, int burnin) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_SIM(ctx);
  return guarded([&] {
    ctx->grain->endIntegration(burnin != 0);
    return SC_OK;
  });
}

int sc_num_fields(sc_context *ctx, int *numFields) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_SIM(ctx);
  SC_CHECK_PTR(numFields);
  *numFields = ctx->grain->numFields();
  return SC_OK;
}

int sc_accelerator_step(sc_context *ctx, double *x, const double *g) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_SIM(ctx);
  SC_CHECK_PTR(x);
  SC_CHECK_PTR(g);
  return guarded([&] {
    ctx->grain->acceleratorStepOn(x, g);
    return SC_OK;
  });
}

int sc_poisson_solve(int n, double h, int numGrain, int enforce, const double *current,
                     double *out) {
  return sc_poisson_solve_ex(n, h, numGrain, enforce, 0.0, current, out);
}

int sc_poisson_solve_ex(int n, double h, int numGrain, int enforce, double maxError,
                        const double *current, double *out) {
  SC_CHECK_PTR(current);
  SC_CHECK_PTR(out);
  if (n < 2 || !(h > 0.0) || numGrain < 1)
    return fail(SC_ERR_ARGUMENT, "need n >= 2, h > 0 and num_grain >= 1");
  if (!(maxError >= 0.0)) return fail(SC_ERR_ARGUMENT, "max_error must be >= 0");
  return guarded([&] {
    scamd::poissonSolve(n, h, numGrain, enforce != 0, current, out, maxError);
    return SC_OK;
  });
}

int sc_poisson_kernel(int n, double h, double maxError, double *out, int *numFilters) {
  SC_CHECK_PTR(out);
  if (n < 1 || !(h > 0.0)) return fail(SC_ERR_ARGUMENT, "need n >= 1 and h > 0");
  if (!(maxError >= 0.0)) return fail(SC_ERR_ARGUMENT, "max_error must be >= 0");
  return guarded([&] {
    std::vector<double> g;
    const int k = scamd::poissonKernel(n, h, g, maxError);
    std::copy(g.begin(), g.end(), out);
    if (numFilters) *numFilters = k;
    return SC_OK;
  });
}

int sc_riccati_steps(int n, const double *in, double *out, int precision) {
  SC_CHECK_PTR(in);
  SC_CHECK_PTR(out);
  if (n < 0) return fail(SC_ERR_ARGUMENT, "n must be non-negative");
  if (n == 0) return SC_OK;
  return guarded([&] {
    checkPrecision(precision);
    double *din = nullptr, *dout = nullptr;
    hipCheck(hipMalloc(reinterpret_cast<void **>(&din), size_t(n) * 7 * sizeof(double)), "alloc");
    hipCheck(hipMalloc(reinterpret_cast<void **>(&dout), size_t(n) * 2 * sizeof(double)), "alloc");
    hipCheck(hipMemcpy(din, in, size_t(n) * 7 * sizeof(double), hipMemcpyHostToDevice), "copy");
    riccatiStepsKernel<<<(n + 255) / 256, 256>>>(n, din, dout, precision);
    hipCheck(hipGetLastError(), "riccati steps");
    hipCheck(hipMemcpy(out, dout, size_t(n) * 2 * sizeof(double), hipMemcpyDeviceToHost), "copy");
    (void)hipFree(din);
    (void)hipFree(dout);
    return SC_OK;
  });
}

int sc_num_ozaki(double t, double cutoff) {
  if (!(t > 0.0) || !(cutoff > 0.0)) return fail(SC_ERR_ARGUMENT, "need t > 0, cutoff > 0") * -1;
  return scamd::numOzakiEnergies(t, cutoff);
}
int sc_ozaki(int n, double *poles, double *residues) {
  SC_CHECK_PTR(poles);
  SC_CHECK_PTR(residues);
  if (n < 1) return fail(SC_ERR_ARGUMENT, "need at least one pole");
  std::vector<double> p, r;
  scamd::ozakiPoles(n, p, r);
  std::memcpy(poles, p.data(), n * sizeof(double));
  std::memcpy(residues, r.data(), n * sizeof(double));
  return SC_OK;
}
double sc_bulk_gap(double t) { return scamd::bulkGapSWave(t); }

int sc_describe(const char *config_json, int angle, sc_dims *out, signed char *labels) {
  SC_CHECK_PTR(config_json);
  SC_CHECK_PTR(out);
  return guarded([&] {
    const auto s = scamd::buildSetup(scamd::parseConfig(config_json));
    out->n0 = s->N0;
    out->n1 = s->N1;
    out->num_energies = s->Ne;
    out->num_angles = s->Na;
    out->num_components = s->Nc;
    out->num_grain_sites = s->numGrain;
    out->num_in_gamma = s->nInG;
    out->num_out_gamma = s->nOutG;
    out->num_in_gammatilde = s->nInGt;
    out->num_out_gammatilde = s->nOutGt;
    out->traj_steps = s->trajSteps;
    out->h = s->h;
    out->grain_fraction = s->grainFraction;
    if (labels) {
      if (angle < 0 || angle >= s->Na) throw std::invalid_argument("angle out of range");
      const size_t n1 = size_t(s->N1) * s->N1;
      std::memcpy(labels, s->labels.data() + size_t(angle) * n1, n1);
    }
    return SC_OK;
  });
}

int sc_get_domain(sc_context *ctx, signed char *domain) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_PTR(domain);
  const scamd::Setup &s = ctx->grain->setup();
  for (size_t i = 0; i < s.domain.size(); ++i) domain[i] = static_cast<signed char>(s.domain[i]);
  return SC_OK;
}

int sc_get_labels(sc_context *ctx, int angle, signed char *labels) {
  SC_CHECK_CTX(ctx);
  SC_CHECK_PTR(labels);
  const scamd::Setup &s = ctx->grain->setup();
  if (angle < 0 || angle >= s.Na) return fail(SC_ERR_ARGUMENT, "angle out of range");
  const size_t n1 = size_t(s.N1) * s.N1;
  std::memcpy(labels, s.labels.data() + size_t(angle) * n1, n1);
  return SC_OK;
}

// ----- batches -----
int sc_batch_create(const char *const *config_json, int count, sc_batch **out) {
  SC_CHECK_PTR(config_json);
  SC_CHECK_PTR(out);
  *out = nullptr;
  if (count < 1) return fail(SC_ERR_ARGUMENT, "a batch needs at least one config");
  return guarded([&] {
    std::vector<std::unique_ptr<scamd::Setup>> setups;
    for (int i = 0; i < count; ++i) {
      if (!config_json[i]) throw std::invalid_argument("null config in the batch");
      setups.push_back(scamd::buildSetup(scamd::parseConfig(config_json[i])));
    }
    auto b = std::make_unique<sc_batch>();
    b->batch = std::make_unique<scamd::Batch>(std::move(setups));
    for (int i = 0; i < count; ++i) {
      auto v = std::make_unique<sc_context>();
      v->grain = &b->batch->grain(i);
      v->borrowed = true;
      b->views.push_back(std::move(v));
    }
    *out = b.release();
    return SC_OK;
  });
}
int sc_batch_destroy(sc_batch *b) {
  if (!b) return SC_OK;
  return guarded([&] {
    delete b;
    return SC_OK;
  });
}
int sc_batch_size(sc_batch *b, int *count) {
  SC_CHECK_PTR(b);
  SC_CHECK_PTR(count);
  *count = b->batch->size();
  return SC_OK;
}
int sc_batch_grain(sc_batch *b, int index, sc_context **out) {
  SC_CHECK_PTR(b);
  SC_CHECK_PTR(out);
  if (index < 0 || index >= b->batch->size()) return fail(SC_ERR_ARGUMENT, "grain index out of range");
  *out = b->views[size_t(index)].get();
  return SC_OK;
}
int sc_batch_compute(sc_batch *b, int precision) {
  SC_CHECK_PTR(b);
  return guarded([&] {
    checkPrecision(precision);
    b->batch->compute(precision);
    return SC_OK;
  });
}
int sc_batch_burnin(sc_batch *b, int precision) {
  SC_CHECK_PTR(b);
  return guarded([&] {
    checkPrecision(precision);
    b->batch->burnin(precision);
    return SC_OK;
  });
}
int sc_batch_run(sc_batch *b, int nb, int nmin, int nmax, int precision, int *converged) {
  SC_CHECK_PTR(b);
  return guarded([&] {
    if (precision != SC_PRECISION_MIXED_REFINE) checkPrecision(precision);
    const scamd::Config &c = b->batch->grain(0).setup().cfg;
    std::vector<int> conv;
    b->batch->run(nb < 0 ? c.burnin : nb, nmin < 0 ? c.itMin : nmin, nmax < 0 ? c.itMax : nmax,
                  precision, conv);
    if (converged)
      for (size_t i = 0; i < conv.size(); ++i) converged[i] = conv[i];
    return SC_OK;
  });
}
int sc_batch_newton_krylov(sc_batch *b, int maxNewton, int krylovDim, double tol,
                           int jvpPrecision, const int *grains, sc_newton_report *reports) {
  SC_CHECK_PTR(b);
  return guarded([&] {
    if (maxNewton < 0) throw std::invalid_argument("max_newton must be >= 0");
    if (krylovDim < 1 || krylovDim > 63)
      throw std::invalid_argument("krylov_dim must be in [1, 63]");
    if (jvpPrecision != SC_PRECISION_FP64 && jvpPrecision != SC_PRECISION_MIXED &&
        jvpPrecision != SC_JVP_TANGENT)
      throw std::invalid_argument(
          "jvp_precision must be SC_PRECISION_FP64, SC_PRECISION_MIXED or SC_JVP_TANGENT");
    if (!(tol > 0.0)) throw std::invalid_argument("tol must be positive");
    const int n = b->batch->size();
    if (reports)
      for (int i = 0; i < n; ++i)
        if (reports[i].history_capacity > 0 && !reports[i].history)
          throw std::invalid_argument("history is null with a positive capacity");
    std::vector<int> which;
    if (grains) which.assign(grains, grains + n);
    const std::vector<scamd::NewtonReport> rs = b->batch->newtonKrylov(
        maxNewton, krylovDim, tol,
        jvpPrecision == SC_PRECISION_MIXED
            ? scamd::PREC_MIXED
            : (jvpPrecision == SC_JVP_TANGENT ? scamd::JVP_TANGENT : scamd::PREC_FP64),
        which);
    if (reports)
      for (int i = 0; i < n; ++i) {
        const scamd::NewtonReport &r = rs[size_t(i)];
        sc_newton_report &o = reports[i];
        o.newton_iterations = r.newtonIterations;
        o.krylov_iterations = r.krylovIterations;
        o.map_evaluations = r.mapEvaluations;
        o.converged = r.converged ? 1 : 0;
        o.residual = r.residual;
        const int h = std::min(static_cast<int>(r.history.size()), std::max(o.history_capacity, 0));
        for (int q = 0; q < h; ++q) o.history[q] = r.history[static_cast<size_t>(q)];
        o.history_length = h;
      }
    return SC_OK;
  });
}
int sc_batch_set_sweep_shape(sc_batch *b, int precision, int el, int waves, int slots) {
  SC_CHECK_PTR(b);
  return guarded([&] {
    checkPrecision(precision);
    b->batch->setSweepShape(precision, el, waves, slots);
    return SC_OK;
  });
}
int sc_batch_set_graphs(sc_batch *b, int enable) {
  SC_CHECK_PTR(b);
  return guarded([&] {
    for (int i = 0; i < b->batch->size(); ++i) b->batch->grain(i).setGraphs(enable != 0);
    return SC_OK;
  });
}
int sc_batch_get_sweep_shape(sc_batch *b, int precision, int *el, int *waves, int *slots) {
  SC_CHECK_PTR(b);
  SC_CHECK_PTR(el);
  SC_CHECK_PTR(waves);
  SC_CHECK_PTR(slots);
  return guarded([&] {
    checkPrecision(precision);
    b->batch->sweepShape(precision, *el, *waves, *slots);
    return SC_OK;
  });
}
int sc_batch_autotune(sc_batch *b, int precision) {
  SC_CHECK_PTR(b);
  return guarded([&] {
    checkPrecision(precision);
    b->batch->autotune(precision);
    return SC_OK;
  });
}
int sc_batch_timing(sc_batch *b, double *sweep_ms, double *iteration_ms,
                    long long *traj_steps) {
  SC_CHECK_PTR(b);
  if (sweep_ms) *sweep_ms = b->batch->lastSweepMs();
  if (iteration_ms) *iteration_ms = b->batch->lastIterationMs();
  if (traj_steps) *traj_steps = b->batch->lastTrajSteps();
  return SC_OK;
}

} // extern "C"
