/*
 * superconga_amd.h -- C ABI of the MI355X self-consistency path.
 *
 * A drop-in for SuperConga's self-consistency hot path: the Riccati solver,
 * the specular boundary condition, the angle/energy reductions into order
 * parameter, current and free energy, the vector-potential solve and the
 * accelerators, all on one GPU.  The entry points replace the reference's
 * C++ module interfaces (file:line citations below are into the reference
 * tree, superconga 1.1.0).  Host buffers use the reference's GridData layout:
 * row-major y*N + x per field, separate real and imaginary planes.
 *
 * Every function returns SC_OK or an error code; sc_last_error() returns a
 * thread-local message for the last failure.  Handles are independent; one
 * handle must not be used from two threads at once.
 */
#ifndef SUPERCONGA_AMD_H_
#define SUPERCONGA_AMD_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SC_OK 0
#define SC_ERR_ARGUMENT 1
#define SC_ERR_CONFIG 2
#define SC_ERR_DEVICE 3
#define SC_ERR_STATE 4

/* Arithmetic of the Riccati stepping. */
#define SC_PRECISION_FP64 0  /* everything in FP64 */
#define SC_PRECISION_MIXED 1 /* FP32 stepping, FP64 energy sums and reductions */
#define SC_PRECISION_MIXED_REFINE 2 /* sc_run, sc_batch_run: MIXED to the FP32 floor, then FP64 */
#define SC_JVP_TANGENT 3 /* sc_newton_krylov_ex, sc_batch_newton_krylov: exact products */

typedef struct sc_context sc_context;
typedef struct sc_batch sc_batch;

typedef struct {
  int n0;               /* reference lattice side (Parameters::getGridResolutionBase) */
  int n1;               /* rotated-frame lattice side (getGridResolutionCoherence) */
  int num_energies;     /* Ozaki poles */
  int num_angles;       /* Fermi-surface angles */
  int num_components;   /* order-parameter components */
  int num_grain_sites;  /* GeometryGroup::getNumGrainLatticeSites */
  long long num_in_gamma, num_out_gamma, num_in_gammatilde, num_out_gammatilde;
  long long traj_steps; /* trajectory steps per iteration (2 Ne sum of inside sites) */
  double h;             /* lattice spacing in xi0 */
  double grain_fraction;
} sc_dims;

typedef struct {
  double residual; /* Context::getResidual (Context.h:462-483) */
  double residual_boundary, residual_order_parameter, residual_current,
      residual_free_energy, residual_vector_potential;
  double free_energy, free_energy_per_area, magnetic_free_energy,
      magnetic_free_energy_per_area; /* ComputeFreeEnergy getters */
  int iterations;
  int converged; /* Context::isPartiallyConverged (Context.h:485-497) */
  double last_sweep_ms;     /* Riccati sweep kernel time of the last iteration */
  double last_iteration_ms; /* whole iteration */
  int mixed_iterations;     /* FP32-stepping iterations of the last sc_run */
} sc_results;

const char *sc_last_error(void);
int sc_version(void);

/* Builds a context from the text of a simulation_config.json (the file the
 * reference's simulation binary reads, src/simulation.cu:64-463): parameters,
 * geometry, Ozaki energies, order parameter with its initial guess,
 * accelerator and compute objects.  Replaces Context<T>::initialize
 * (Context.h:285-375). */
int sc_create(const char *config_json, sc_context **out);
int sc_destroy(sc_context *ctx);
int sc_get_dims(sc_context *ctx, sc_dims *out);
/* Ozaki energies (imaginary parts) and residues, IntegrationIteratorOzaki.h:165-200. */
int sc_get_energies(sc_context *ctx, double *energies, double *residues);

/* Order parameter component `comp` (OrderParameter::getComponentGrid,
 * OrderParameter.h:214-230).  which = 0: the iterate x; which = 1: the
 * gap-equation result g(x) of the last compute (ComputeOrderParameter::
 * getResult).  Setting masks the sites outside the grain like
 * OrderParameter::initialize (OrderParameter.h:180-199). */
int sc_set_order_parameter(sc_context *ctx, int comp, const double *re,
                           const double *im);
int sc_get_order_parameter(sc_context *ctx, int which, int comp, double *re,
                           double *im);
/* Induced vector potential in units of A0 (unscaled).  which = 0: iterate,
 * 1: last Poisson solution.  The reference stores it scaled by lambda^2
 * (src/simulation.cu:369-376). */
int sc_set_vector_potential(sc_context *ctx, const double *ax, const double *ay);
int sc_get_vector_potential(sc_context *ctx, int which, double *ax, double *ay);
/* Current density of the last compute (ComputeCurrent::getResult). */
int sc_get_current(sc_context *ctx, double *jx, double *jy);
/* Free-energy density Re I(r) summed over angles: the ComputeFreeEnergy grid
 * (ComputeFreeEnergy.h:281-300) before its factor T. */
int sc_get_free_energy_density(sc_context *ctx, double *out);
/* Sparse boundary storage (BoundaryStorage.h:158-169, layout
 * (crossing * Ne + energy)).  in_out: 0 in, 1 out; dir: 0 gamma, 1 gamma~. */
int sc_get_boundary(sc_context *ctx, int in_out, int dir, double *re, double *im);
int sc_set_boundary(sc_context *ctx, int dir, const double *re, const double *im);

/* One self-consistency iteration, Context<T>::compute (Context.h:391-447):
 * accelerator step (from the second call on), Riccati sweeps over all angles
 * and energies fused with the Green-function sums, angle integration, gap
 * equation, current, free energy, specular boundary condition and, with a
 * finite penetration depth, the vector-potential solve. */
int sc_compute(sc_context *ctx, int precision);
/* One burn-in iteration, Context<T>::burnIn (Context.h:377-389). */
int sc_burnin(sc_context *ctx, int precision);
/* runSimulation (run_simulation.h:124-187) with the limits of the config
 * when the arguments are negative; *converged receives the final status.
 * With SC_PRECISION_MIXED_REFINE the iterations step in FP32 until the
 * residual reaches the FP32 floor and then in FP64; convergence is only
 * declared on an FP64 iteration. */
int sc_run(sc_context *ctx, int num_burnin, int num_min, int num_max,
           int precision, int *converged);
int sc_get_results(sc_context *ctx, sc_results *out);

/* ----- Keldysh spectroscopy (src/postprocess.cu) ----- */

/* A postprocess context: the simulation config's physics and geometry, the
 * postprocess config's energies (spectroscopy.energy_min/max, num_energies,
 * energy_broadening), angles, criterion, norm, energies per block and
 * iteration limits (src/postprocess.cu:64-330).  The order parameter starts
 * at zero and the vector potential at zero: set them with
 * sc_set_order_parameter / sc_set_vector_potential (the reference reads
 * order_parameter.csv and vector_potential.csv of misc.load_path).  Such a
 * context runs only the sc_ldos_* iteration and the sweep-shape calls. */
int sc_ldos_create(const char *simulation_config_json, const char *postprocess_config_json,
                   sc_context **out);
/* Context<T>::compute / burnIn of the postprocess context: one iteration of
 * IntegrationIteratorKeldysh::compute (IntegrationIteratorKeldysh.h:235-388)
 * with ComputeLDOS (ComputeLDOS.h:110-192) on the current energy block. */
int sc_ldos_compute(sc_context *ctx, int precision, int burnin);
/* runPostprocess (run_simulation.h:44-108): for every energy block a burn-in
 * of num_burnin iterations (-1: until the boundary converges) and then
 * compute iterations between num_min and num_max (-1: no maximum) until the
 * block converges; stops at the first block that does not.  Pass the
 * postprocess config's num_iterations_* to reproduce the reference. */
int sc_ldos_run(sc_context *ctx, int num_burnin, int num_min, int num_max, int precision,
                int *converged);
typedef struct {
  int num_energies, energies_per_block;
  int block_start, block_count; /* current energy block */
  int iterations;
  int angle_chunk;              /* angles whose spectral weights are held at once */
  int partially_converged;      /* Context::isPartiallyConverged */
  int converged;                /* Context::isConverged: every block done */
  double residual;              /* max of the two below (Context::getResidual) */
  double residual_boundary, residual_ldos;
  double last_sweep_ms, last_iteration_ms;
  long long traj_steps;         /* per iteration of the current block */
} sc_ldos_status;
int sc_ldos_get_status(sc_context *ctx, sc_ldos_status *out);
/* The LDOS of every energy, ComputeLDOS::getResult: [num_energies][N0][N0]. */
int sc_ldos_get(sc_context *ctx, double *out);
/* Angles swept per chunk of spectral weights (0: as many as half the free
 * device memory holds, at most 64 GiB). */
int sc_ldos_set_angle_chunk(sc_context *ctx, int angles);

/* Outcome of sc_newton_krylov. */
typedef struct {
  int newton_iterations;  /* accepted Newton steps */
  int krylov_iterations;  /* GMRES directions over all steps */
  int map_evaluations;    /* plain iterations G evaluated */
  int converged;          /* 1 when |G(u) - u|_W fell below tol */
  double residual;        /* final |G(u) - u|_W */
  double *history;        /* optional, caller-owned: |G(u) - u|_W per step */
  int history_capacity;   /* entries history can hold */
  int history_length;     /* entries written */
} sc_newton_report;

/* Jacobian-free Newton-Krylov on the fixed point u = G(u) of one plain
 * iteration (no reference counterpart: the reference only has the fixed-point
 * accelerators of accelerators/, Context.h:391-447).  u stacks the order
 * parameter, the induced vector potential when the penetration depth is
 * finite, and the boundary in-values of both coherence functions; G is one
 * FP64 sweep + angle integration + gap equation + boundary condition +
 * Poisson solve.  Each Newton step solves J d = -(G(u) - u) by GMRES
 * restarted every krylov_dim (1..63) directions, at most 4 cycles, with
 * finite-difference products, in the inner product weighted per block by
 * 1 / |u_block|^2,
 * followed by a backtracking line search; the forcing term follows
 * Eisenstat-Walker.  The solve ends at tol, at max_newton steps, when the
 * line search finds no descent, or after two accepted steps in a row that
 * each keep more than 0.9 of the residual (the rounding floor of the map:
 * converged then stays 0).  On return the state holds the last accepted u
 * and the accelerator restarts; a following sc_compute reports the
 * reference's residuals for it. */
int sc_newton_krylov(sc_context *ctx, int max_newton, int krylov_dim, double tol,
                     sc_newton_report *report);
/* sc_newton_krylov with the precision of the Jacobian-vector products:
 * SC_PRECISION_FP64 (as sc_newton_krylov), SC_PRECISION_MIXED (both ends
 * of each finite difference on the map with FP32 stepping and FP64 sums, at
 * a step of 3e-4 (1 + |u|) in the weighted norm; the Newton residuals stay
 * FP64, so the solution is the FP64 fixed point; a Newton step whose line
 * search finds no descent hands the rest of the solve to FP64 products,
 * see sc_set_newton_options) or SC_JVP_TANGENT (exact
 * FP64 products J v by the tangent-linear sweep, as sc_jacobian_vector
 * forms them, with the boundary condition's and the screened update's
 * tangents).  krylov_dim 1..63. */
int sc_newton_krylov_ex(sc_context *ctx, int max_newton, int krylov_dim, double tol,
                        int jvp_precision, sc_newton_report *report);

/* The free energy one iteration reports from the current state (the
 * reference's ComputeFreeEnergy.h:281-414 on the Green functions of one
 * sweep) and its gradient with respect to the order parameter by a discrete
 * adjoint of that iteration: the transposes of the back-rotation and of the
 * extensions, a reverse sweep of both Riccati recurrences, the transpose of
 * the rotated order parameter.  The vector potential and the boundary
 * in-values are held fixed.  grad holds num_components x 2 x N0^2 values:
 * dF/dRe Delta_c, then dF/dIm Delta_c, per component in order.  The state is
 * left as it was.  No counterpart in the reference: it serves gradient
 * studies (torch.autograd in the Python frontend).  Its FP64 atomic
 * scatters make two calls agree to rounding, not bit for bit. */
int sc_free_energy_gradient(sc_context *ctx, double *grad, double *free_energy);

/* sc_free_energy_gradient with, when grad_a is not NULL, also the gradient
 * with respect to the induced vector potential: grad_a holds 2 x N0^2
 * values, dF/dA_x then dF/dA_y, per unit of A as sc_set_vector_potential
 * takes it.  A enters F through the Doppler shift of every Riccati step
 * (the reverse sweep's adjoint of the energy shift, transposed through the
 * bilinear taps of A) and through the magnetic term; without a penetration
 * depth grad_a is zero.  No counterpart in the reference, like
 * sc_free_energy_gradient. */
int sc_free_energy_gradient_ex(sc_context *ctx, double *grad, double *grad_a,
                               double *free_energy);

/* The self-consistency residual of one plain FP64 iteration u -> G(u) (the
 * gap equation, the current and its vector potential without accelerator;
 * the boundary in-values held fixed),
 *   R = 1/2 sum over inside sites of |Delta'_c - Delta_c|^2 (all components)
 *       + 1/2 sum over the grid of |A' - A|^2 (A as sc_set_vector_potential
 *       takes it),
 * and its gradient (J - I)^T (G(u) - u): grad as sc_free_energy_gradient's,
 * grad_a (may be NULL; 2 x N0^2, zero without a penetration depth) in A.
 * The same reverse sweep as the free-energy gradient, seeded through the
 * Green functions' derivatives from the gap equation and the current.  The
 * state is left as it was.  No counterpart in the reference. */
int sc_residual_gradient(sc_context *ctx, double *grad, double *grad_a, double *residual);

/* The Jacobian-vector product J v of one plain FP64 iteration u -> G(u)
 * (the map of sc_residual_gradient: gap equation, current and vector
 * potential, boundary in-values held fixed) at the current state, exactly,
 * by the tangent-linear sweep: both Riccati recurrences with their
 * tangents, the Green functions' tangents, then the linear stages (the
 * extensions and site fields, the back-rotation and angle integration, the
 * Poisson solve) on the tangents.  v and jv hold num_components x 2 x N0^2
 * values (dRe Delta_c, dIm Delta_c) then 2 x N0^2 (dA_x, dA_y, in the units
 * sc_set_vector_potential takes; ignored, and zero in jv, without a
 * penetration depth).  The state is left as it was.  No counterpart in the
 * reference. */
int sc_jacobian_vector(sc_context *ctx, const double *v, double *jv);

/* How the reverse sweeps of sc_free_energy_gradient(_ex) and
 * sc_residual_gradient keep the coherence functions they run back over:
 * k > 0 keeps a checkpoint every k sites and integrates each segment again
 * when the reverse passes reach it (twice the forward steps, scratch of
 * ceil(L / k) checkpoints and one segment instead of L sites for the
 * longest trajectory L); -1 (the default) takes k = ceil(sqrt(L)); 0 keeps
 * every site of every trajectory.  The results are the same.  No
 * counterpart in the reference. */
int sc_set_adjoint_checkpoints(sc_context *ctx, int k);

/* Newton-Krylov on this grain (sc_newton_krylov(_ex), and
 * sc_batch_newton_krylov for the batch's grains): GMRES cycles per Newton
 * step (restarts >= 1, default 4), the cap of the Eisenstat-Walker forcing
 * term (0 < eta_max < 1, default 0.3), the linear residual a Newton step
 * asks for relative to the nonlinear one, and with SC_PRECISION_MIXED
 * products whether a line search without descent hands the solve over to
 * FP64 products (fp64_fallback != 0, the default) or ends it, and whether
 * each Newton step loses its component along the generator of the order
 * parameter's global phase rotation (fix_phase != 0; default 0), a
 * direction in which J - I is singular because every rotation of a fixed
 * point is one.  No counterpart in the reference. */
int sc_set_newton_options(sc_context *ctx, int restarts, double eta_max, int fp64_fallback,
                          int fix_phase);

/* Individual stages. */
/* RiccatiSolver::computeCoherenceFunctions for every angle (RiccatiSolverConfined.h:513-671)
 * fused with ComputeProperty::computeGreensAngle; burnin = 1 only updates the
 * boundary out-values (computeCoherenceFunctionsBurnIn). */
int sc_riccati_sweep(sc_context *ctx, int precision, int burnin);
/* BoundaryConditionSpecular::computeBoundaryCondition (BoundaryConditionSpecular.h:361-486). */
int sc_boundary_condition(sc_context *ctx, double *residual);

/* One iteration angle by angle, the way IntegrationIteratorOzaki::compute
 * drives the RiccatiSolver and the ComputeProperty hooks
 * (IntegrationIteratorOzaki.h:300-420, ComputeProperty.h:27-103):
 *   sc_integration_begin(ctx, burnin)
 *       Context<T>::compute up to the angle loop (Context.h:391-447: the
 *       accelerator step from the second iteration on) and
 *       ComputeProperty::initializeIteration;
 *   for consecutive ranges [angle_begin, angle_end) covering every angle:
 *     sc_sweep_angles(ctx, precision, burnin, angle_begin, angle_end)
 *       RiccatiSolver::computeCoherenceFunctions(BurnIn) of those angles
 *       (RiccatiSolverConfined.h:513-671) with computeGreensAngle;
 *     sc_integrate_angles(ctx, angle_begin, angle_end)
 *       ComputeProperty::integrateGreensAngle of those angles
 *       (ComputeOrderParameter.h:174-210, ComputeCurrent.h:142-173,
 *       ComputeFreeEnergy.h:260-279); ranges in increasing order, none in
 *       a burn-in;
 *   sc_integration_end(ctx, burnin)
 *       ComputeProperty::computePropertyFromGreens, the boundary condition,
 *       the vector-potential solve and the residuals (the rest of
 *       Context<T>::compute / burnIn).
 * The result equals sc_compute / sc_burnin bit for bit.  Calls out of this
 * order return SC_ERR_STATE-like errors (SC_ERR_DEVICE with the message). */
int sc_integration_begin(sc_context *ctx, int burnin);
int sc_sweep_angles(sc_context *ctx, int precision, int burnin, int angle_begin, int angle_end);
int sc_integrate_angles(sc_context *ctx, int angle_begin, int angle_end);
int sc_integration_end(sc_context *ctx, int burnin);

/* Accelerator<T>::compute (accelerators/Accelerator.h:25-71) on
 * caller-supplied fields: x is the iterate and g the map's result, each the
 * concatenated field of accelerators/utils.h:58-186 -- Re and Im of every
 * order-parameter component, then lambda^2 Ax and lambda^2 Ay when the
 * context has a finite penetration depth -- of num_fields x N0^2 values
 * (sc_num_fields).  x receives the accelerated iterate.  The accelerator
 * (name, rank, step size ... of the config) and its history are the
 * context's own, so consecutive calls continue one acceleration; the
 * context's order parameter and vector potential are overwritten. */
int sc_num_fields(sc_context *ctx, int *num_fields);
int sc_accelerator_step(sc_context *ctx, double *x, const double *g);
/* Lane layout of the sweep for one precision (no reference counterpart: the
 * reference's shape rule is getKernelDimensions, GridFunctions_GPU.h:1121-1189):
 * Ozaki energies per lane, the register budget the kernel variant is compiled
 * for (wavefronts per SIMD) and persistent wavefronts per compute unit.
 * Compiled variants (energies, waves): (1,4) (2,4) (3,3) (4,2) (4,3) (4,4)
 * (5,2) (6,2) (7,1) (7,2) (7,4) (8,1) (8,2) (8,4) (14,1) (14,2) (16,1)
 * (16,2); for spectroscopy contexts (2,4) (4,2) (4,3) (4,4) (8,1) (8,2)
 * (8,4) (16,1). */
/* The device work after each iteration's sweep (back-rotation and angle
 * integration, gap equation, current, free energy and residual reductions,
 * boundary condition, Poisson solve) is captured into a hipGraph on its
 * second run and replayed after that, while the reductions it was captured
 * with keep their offsets (a change re-captures it).  The graph holds only
 * kernel and copy nodes.  Results are bit for bit those of the eager
 * launches.  On by default; enable = 0 launches every stage eagerly.  No
 * counterpart in the reference, which launches each stage per iteration
 * (Context.h:391-447). */
int sc_set_graphs(sc_context *ctx, int enable);
/* Switch the screened vector-potential update (config extension
 * accelerator.vector_potential_screening) on with the given superfluid
 * density (> 0) or off (0) between iterations; the accelerator restarts.
 * The fixed points are unchanged.  No counterpart in the reference. */
int sc_set_screening(sc_context *ctx, double superfluid_density);
int sc_set_sweep_shape(sc_context *ctx, int precision, int energies_per_lane,
                       int waves_per_simd, int slots_per_cu);
int sc_get_sweep_shape(sc_context *ctx, int precision, int *energies_per_lane,
                       int *waves_per_simd, int *slots_per_cu);
/* Autotuner: times every variant on the current state and keeps the fastest
 * for `precision`.  The sweeps it runs do not change the iteration state. */
int sc_autotune_sweep(sc_context *ctx, int precision);

/* ComputeVectorPotential with vector_potential_error = 0
 * (ComputeVectorPotential.h:143-201,647-745): A = -h^2 G * (j - <j>_grain).
 * current and out are 2*n*n (x plane then y plane). */
int sc_poisson_solve(int n, double h, int num_grain, int enforce_mean,
                     const double *current, double *out);
/* The same with ComputeVectorPotential's max_error (the config's
 * numerics.vector_potential_error, src/simulation.cu:428-434): > 0 replaces
 * the exact kernel by the eigen-expansion the reference's separable filters
 * sum to (ComputeVectorPotential.h:319-395,686-722), truncated at that
 * relative Frobenius error. */
int sc_poisson_solve_ex(int n, double h, int num_grain, int enforce_mean, double max_error,
                        const double *current, double *out);
/* Largest grid side of sc_poisson_solve_noalloc. */
#define SC_NOALLOC_MAX_GRID 512
/* sc_poisson_solve (exact kernel) without any device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under graph capture, tight serving loops, harnesses
 * that forbid the allocator).  The transform buffers and the staging of
 * current and result live in one static device array of about 34 MB per
 * GPU, held from the load of the library until the process ends; grids
 * above SC_NOALLOC_MAX_GRID sites a side are rejected with SC_ERR_ARGUMENT.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * race.  Runs on the null stream of the current device and returns when out
 * holds the result. */
int sc_poisson_solve_noalloc(int n, double h, int num_grain, int enforce_mean,
                             const double *current, double *out);
/* Host only: the kernel G(|dx|, |dy|) the solve convolves with, n*n values
 * out[dy*n + dx] (ComputeVectorPotential::initializeFilters,
 * ComputeVectorPotential.h:296-417); *num_filters (may be null) receives
 * the number of rank-one filters of the truncation (0 for max_error = 0). */
int sc_poisson_kernel(int n, double h, double max_error, double *out, int *num_filters);

/* Unit-level device arithmetic for closed-form checks: n midpoint steps
 * (RiccatiSolutions.h:65-94).  in = n records of (delta re, im, energy re,
 * im, initial re, im, step); out = n records of (re, im). */
int sc_riccati_steps(int n, const double *in, double *out, int precision);
/* Host-side tables exposed for checks. */
int sc_num_ozaki(double temperature, double cutoff);
int sc_ozaki(int num_poles, double *poles, double *residues);
double sc_bulk_gap(double temperature);
/* Reference-frame domain, 1 inside the grain and 0 outside (N0*N0;
 * GeometryGroup::getDomainGrid, GeometryGroup.h:906-925). */
int sc_get_domain(sc_context *ctx, signed char *domain);
/* Geometry labels of frame `angle` (N1*N1, GeometryGroup::create(-theta)). */
int sc_get_labels(sc_context *ctx, int angle, signed char *labels);
/* Host-only planning (no device needed): the sizes and, if labels is not
 * null, the labels of frame `angle` that sc_create would build for a config. */
int sc_describe(const char *config_json, int angle, sc_dims *out,
                signed char *labels);

/* ----- Batches: many grains of a parameter sweep iterated together -----
 * No counterpart in the reference, whose sweeps are sequential processes
 * warm-started from files (examples/parameter_sweeps/, shell scripts).  Every iteration
 * runs the per-grain stages on one stream and sweeps all grains that are
 * still iterating in one persistent launch over a shared work queue.  The
 * handles returned by sc_batch_grain are borrowed: they work with every
 * per-context accessor above and are destroyed with the batch. */
int sc_batch_create(const char *const *config_json, int count, sc_batch **out);
int sc_batch_destroy(sc_batch *batch);
int sc_batch_size(sc_batch *batch, int *count);
int sc_batch_grain(sc_batch *batch, int index, sc_context **out);
int sc_batch_compute(sc_batch *batch, int precision);
int sc_batch_burnin(sc_batch *batch, int precision);
/* runSimulation per grain (limits of the first config when negative); each
 * grain stops when converged with the minimum reached, or at the maximum.
 * converged (count entries, may be null) receives the final status.  With
 * SC_PRECISION_MIXED_REFINE each grain switches from FP32 stepping to FP64
 * on its own (the rule of sc_run) and converges only on an FP64 iteration;
 * an iteration sweeps the grains of each precision in one launch each. */
int sc_batch_run(sc_batch *batch, int num_burnin, int num_min, int num_max,
                 int precision, int *converged);
/* sc_newton_krylov_ex on every grain of the batch at once: each grain runs
 * its own Newton-Krylov solve (Krylov basis, line search, stopping), and
 * every round evaluates the maps of all grains still solving with one sweep
 * launch per precision (no counterpart in the reference, whose parameter
 * sweeps run one process per point from the shell scripts in
 * examples/parameter_sweeps/).
 * grains (one flag per grain, or null for all) selects the grains to
 * solve; the others are left as they are and their reports zeroed.
 * reports holds one sc_newton_report per grain (or is null).  On return
 * every solved grain's state holds its last accepted u. */
int sc_batch_newton_krylov(sc_batch *batch, int max_newton, int krylov_dim, double tol,
                           int jvp_precision, const int *grains, sc_newton_report *reports);
/* sc_set_graphs for every grain of the batch. */
int sc_batch_set_graphs(sc_batch *batch, int enable);
int sc_batch_set_sweep_shape(sc_batch *batch, int precision, int energies_per_lane,
                             int waves_per_simd, int slots_per_cu);
int sc_batch_get_sweep_shape(sc_batch *batch, int precision, int *energies_per_lane,
                             int *waves_per_simd, int *slots_per_cu);
int sc_batch_autotune(sc_batch *batch, int precision);
/* Last batched iteration: sweep kernel time, whole iteration, and the
 * trajectory steps it took. */
int sc_batch_timing(sc_batch *batch, double *sweep_ms, double *iteration_ms,
                    long long *traj_steps);

#ifdef __cplusplus
}
#endif
#endif
