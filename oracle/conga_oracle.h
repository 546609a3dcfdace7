/*
 * conga_oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a plain C++17/OpenMP FP64
 * restatement of SuperConga's self-consistency path (the reference's CPU
 * backend semantics), written to check the HIP product against.  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load
 * it.  The product never links, loads or calls anything in this directory.
 *
 * Every function cites the reference file:line it restates in
 * conga_oracle.cpp.
 */
#ifndef CONGA_ORACLE_H_
#define CONGA_ORACLE_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OC_MAX_COMPONENTS 4
#define OC_MAX_GEOMETRY 16

/* Geometry component kinds. */
enum { OC_DISC = 0, OC_REGULAR_POLYGON = 1, OC_POLYGON = 2 };
/* Order-parameter symmetries. */
enum { OC_SWAVE = 0, OC_DX2Y2 = 1, OC_DXY = 2, OC_GWAVE = 3 };
/* Gauges of the external vector potential. */
enum { OC_LANDAU = 0, OC_SYMMETRIC = 1, OC_SOLENOID = 2 };
/* Residual norms. */
enum { OC_L1 = 0, OC_L2 = 1, OC_LINF = 2 };
/* Accelerators. */
enum { OC_PICARD = 0, OC_ANDERSON = 1, OC_CONGACC = 2, OC_BB = 3, OC_POLYAK = 4, OC_BROYDEN = 5 };

typedef struct {
  /* physics */
  double temperature;
  double external_flux_quanta;
  double penetration_depth;
  double crystal_axes_rotation; /* radians */
  int gauge;
  int charge_sign;
  int num_components;
  int symmetry[OC_MAX_COMPONENTS];
  double critical_temperature[OC_MAX_COMPONENTS];

  /* geometry.  geometry_internal = 0: user units as in simulation_config.json
   * (bounding box sets the grain width).  geometry_internal = 1: grain_width
   * is given and the components use the internal [-0.5, 0.5] frame:
   * disc (cx, cy, radius), regular polygon (cx, cy, num_edges, rotation
   * [radians], side_length; side_length < 0 means circumradius 0.5). */
  int geometry_internal;
  double grain_width;
  int num_geometry;
  int geometry_type[OC_MAX_GEOMETRY];
  int geometry_add[OC_MAX_GEOMETRY];
  double geometry_param[OC_MAX_GEOMETRY][5];
  int polygon_num_vertices[OC_MAX_GEOMETRY];
  int polygon_vertex_offset[OC_MAX_GEOMETRY];
  const double *polygon_vertices_x;
  const double *polygon_vertices_y;

  /* numerics */
  double convergence_criterion;
  double energy_cutoff;
  int norm;
  int num_fermi_momenta;
  double points_per_coherence_length;
  double vector_potential_error; /* 0: exact Green function; > 0: eigen filters */

  /* compute objects */
  int compute_current;
  int compute_free_energy;

  /* accelerator; non-positive numbers select the reference defaults */
  int accelerator;
  double step_size;
  int rank;
  double drag;
  double step_size_factor;
  double cos_similarity_min;
  double step_size_max;
  double step_size_min;

  /* 1: unwritten coherence-function sites are zero instead of keeping the
   * values of a previous angle (diagnostic; 0 restates the reference). */
  int zero_stale;
  int num_threads; /* <= 0: OpenMP default */
} oc_config;

void *oc_create(const oc_config *config);
void oc_destroy(void *handle);
const char *oc_last_error(void);

/* Scalar queries: "N0", "N1", "num_energies", "num_angles",
 * "num_grain_sites", "num_storage_in_gamma", "num_storage_out_gamma",
 * "num_storage_in_gammatilde", "num_storage_out_gammatilde",
 * "num_iterations", "traj_steps" */
long long oc_get_int(void *handle, const char *key);
/* "residual", "residual_boundary", "residual_order_parameter",
 * "residual_current", "residual_free_energy", "residual_vector_potential",
 * "free_energy", "free_energy_per_area", "magnetic_free_energy",
 * "magnetic_free_energy_per_area", "grain_fraction", "h",
 * "coupling_0".."coupling_3", "delta_bulk_0".."delta_bulk_3",
 * "energy_<i>", "residue_<i>", "system_angle_<k>" */
double oc_get_double(void *handle, const char *key);

/* Arrays: "delta" (context order parameter, Nc*2*N0^2, planar re/im per
 * component), "delta_new" (gap-equation result), "current" (2*N0^2),
 * "vector_potential" (context A scaled by lambda^2, 2*N0^2),
 * "vector_potential_new", "free_energy_grid" (2*N0^2), "domain" (N0^2),
 * "boundary_in_gamma", "boundary_out_gamma", "boundary_in_gammatilde",
 * "boundary_out_gammatilde" (2*count each, planar). Returns the number of
 * doubles written or -1. */
long long oc_get_array(void *handle, const char *name, double *out,
                       long long capacity);
/* Setters (planar re/im). The order parameter is masked outside the grain
 * like OrderParameter::initialize; the vector potential is scaled by
 * lambda^2 like the load path. */
int oc_set_order_parameter(void *handle, int component, const double *re,
                           const double *im);
int oc_set_vector_potential(void *handle, const double *ax, const double *ay);
int oc_set_array(void *handle, const char *name, const double *in,
                 long long count);
int oc_initial_bulk(void *handle, double phase);

/* Geometry of one rotational frame: labels, distances and normals on the
 * N1 x N1 grid for angle index k (frame -systemAngle(k)). */
int oc_get_geometry(void *handle, int angle, signed char *labels,
                    double *distance, double *normal_x, double *normal_y);

int oc_compute(void *handle); /* Context::compute */
int oc_burnin(void *handle);  /* Context::burnIn */
/* runSimulation(burnin, min, max): returns 1 if converged. */
int oc_run(void *handle, int num_burnin, int min_iter, int max_iter);

/* Keldysh spectroscopy, the postprocess context of src/postprocess.cu:
 * energies emin..emax (num of them) with imaginary part `broadening`,
 * iterated in blocks of per_block energies (< 1: all at once) by
 * IntegrationIteratorKeldysh (IntegrationIteratorKeldysh.h:115-392) with
 * ComputeLDOS (ComputeLDOS.h).  Uses the current order parameter and
 * vector potential.  The LDOS is the array "ldos" (num x N0^2). */
int oc_ldos_setup(void *handle, double emin, double emax, int num, double broadening,
                  int per_block);
int oc_ldos_compute(void *handle, int burnin); /* Context::compute / burnIn */
/* runPostprocess (run_simulation.h:44-108): returns 1 if every block converged. */
int oc_ldos_run(void *handle, int num_burnin, int min_iter, int max_iter);

/* Timing helper for the CPU baseline: the per-angle work of one iteration
 * (rotation, energy shift, Riccati sweeps in both directions, Green
 * functions and their integration into the angle sums) for angles
 * [angle_begin, angle_end) with the current state.  Leaves the angle sums
 * partial; returns the number of trajectory steps taken. */
long long oc_sweep_angles(void *handle, int angle_begin, int angle_end);
/* The free energy of one iteration from the current state and its gradient
 * with respect to the order parameter (vector potential and boundary
 * in-values held fixed): grad holds num_components x 2 x N0^2 values,
 * dF/dRe Delta_c then dF/dIm Delta_c.  No counterpart in the reference. */
int oc_free_energy_gradient(void *handle, double *grad, double *free_energy);
/* The same with dF/dA of the induced vector potential (2 N0^2: x plane, then
 * y plane; zero without a penetration depth), A in the units the setters
 * take. */
int oc_free_energy_gradient_a(void *handle, double *grad, double *grad_a, double *free_energy);
/* The self-consistency residual R = 1/2 sum over inside sites of
 * |Delta' - Delta|^2 (every component) + 1/2 sum |A' - A|^2 of one plain
 * iteration with the boundary in-values held fixed, and its gradient in
 * Delta (as above) and in A. */
int oc_residual_gradient(void *handle, double *grad, double *grad_a, double *residual);
/* G(u) of one plain iteration with the boundary in-values held fixed: the
 * gap equation's order parameter (Nc x 2 x N0^2, zero outside the grain) and
 * the vector potential (2 x N0^2, in the setters' units; zero without a
 * penetration depth). */
int oc_plain_map(void *handle, double *delta_new, double *a_new);

/* ----- unit-level restatements (closed-form checks) ----- */
void oc_riccati_step(const double *delta, const double *energy,
                     const double *initial, double step, double *out);
void oc_homogeneous(const double *delta, const double *energy, double *gamma,
                    double *gammatilde);
void oc_greens(const double *gamma, const double *gammatilde, double *g,
               double *f_sum);
int oc_num_ozaki(double temperature, double cutoff);
int oc_ozaki(int num_poles, double *poles, double *residues);
void oc_lattice(double width, double ppxi, int *n0, int *n1,
                double *grain_fraction);
double oc_bulk_gap(double temperature);
double oc_bulk_free_energy(double temperature, double delta);
double oc_yoshida(double temperature, double delta, double cutoff);
/* A bare accelerator on vectors of num_fields * n0^2 values (fields of n0^2
 * values each, the layout of accelerators/utils.h:58-186): oc_accel_step
 * replaces x by the accelerator's next iterate given g = g(x).  The same
 * step functions the iteration uses. */
void *oc_accel_create(int kind, int n0, int num_fields, int rank, double step);
void oc_accel_destroy(void *accel);
int oc_accel_step(void *accel, double *x, const double *g);
/* CongAcc's inertia bisection (CongAcc.h:300-356) for |D|, |M|, D.M. */
double oc_congacc_inertia(double dn, double mn, double dot, double min_cos);

/* ComputeVectorPotential with the exact Green function. */
int oc_vector_potential(int n, double h, int num_grain, int enforce_mean,
                        const double *current, double *out);
/* ComputeVectorPotential(max_error, analytical): max_error > 0 truncates the
 * eigen-expansion of the kernel (ComputeVectorPotential.h:319-395);
 * analytical = 0 takes the plain 2 ln r matrix elements (:281-293). */
int oc_vector_potential_ex(int n, double h, int num_grain, int enforce_mean,
                           double max_error, int analytical, const double *current,
                           double *out);
/* The kernel G(|dx|, |dy|) as out[dy * n + dx]; returns the number of
 * rank-one filters (0 for max_error = 0). */
int oc_poisson_kernel(int n, double h, double max_error, int analytical, double *out);

#ifdef __cplusplus
}
#endif
#endif
