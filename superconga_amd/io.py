"""The reference's on-disk formats, so that its Python frontend and plotters
work unchanged on the files this package writes.

* CSV grids (io/CSVGridWriter.h:30-242): header ``x_idx,y_idx,<titles>``,
  complex titles split into ``<t>_re,<t>_im``, one row per site in row-major
  order (x fastest), floats in C++ ``std::scientific`` with 8 digits,
  integer grids as integers.
* CSV reading (io/CSVReader.h:30-191): columns by header name, ``x_idx`` and
  ``y_idx`` skipped.
* ``simulation_convergence.csv`` (io/CSVSimulationIterationWriter.h:20-101).
* ``simulation_results.json`` and ``internal_parameters.json``
  (io/SimulationWriter.h:50-130).
* ``load_path`` warm starts (src/simulation.cu:296-375, io/csv.h:110-397):
  ``order_parameter.csv`` (columns ``<type>_re/_im``, next to a
  ``domain.csv``) and ``vector_potential.csv`` (``a_x``, ``a_y``).

:func:`simulate` is the in-process counterpart of the reference's
``simulation`` binary (src/simulation.cu:64-530), :func:`postprocess` of its
``postprocess`` binary (src/postprocess.cu:64-472): ``ldos.csv`` (header
``x_idx,y_idx,e_idx,ldos``, one row per site and energy, CSVGridWriter with
one title for many fields), ``postprocess_config.json`` and
``postprocess_convergence.csv`` (io/CSVPostprocessIterationWriter.h).
"""
import json
import math
import os

import numpy as np

# order-parameter component names in the order the reference adds them
# (src/simulation.cu:201-213)
COMPONENT_ORDER = ("dx2-y2", "dxy", "g", "s")
GRAIN_SIZE_MARGIN = 3  # defines.h


def _fmt(v):
    """One value the way std::scientific with precision 8 prints it."""
    s = "%.8e" % v
    return s


# ----------------------------------------------------------------- reading
def read_csv(path):
    """Columns of a reference CSV file as float arrays (CSVReader::read)."""
    with open(path) as f:
        header = f.readline().strip().split(",")
        data = np.loadtxt(f, delimiter=",", ndmin=2) if header else None
    out = {}
    if data is None or data.size == 0:
        return {name: np.empty(0) for name in header if name not in ("x_idx", "y_idx")}
    for i, name in enumerate(header):
        if name in ("x_idx", "y_idx"):
            continue
        out[name] = data[:, i]
    return out


def _square(values, n, what):
    if values.size != n * n:
        side = int(round(math.sqrt(values.size)))
        if side * side != values.size:
            raise ValueError("%s: the data is not square" % what)
        raise NotImplementedError(
            "%s: resampling a %dx%d grid to %dx%d is not supported; load data "
            "from a run on the same lattice" % (what, side, side, n, n))
    return values.reshape(n, n)


def component_types(config):
    """Order-parameter component names of a config in the reference's order."""
    op = config["physics"]["order_parameter"]
    return [name for name in COMPONENT_ORDER if name in op]


def load_state(sim, config, load_path):
    """Warm start from a directory written by the reference or by
    :func:`save_state` (src/simulation.cu:296-375)."""
    n = sim.N0
    op_path = os.path.join(load_path, "order_parameter.csv")
    dom_path = os.path.join(load_path, "domain.csv")
    a_path = os.path.join(load_path, "vector_potential.csv")
    for p in (op_path, dom_path, a_path):
        if not os.path.exists(p):
            raise FileNotFoundError("-- ERROR! Could not open file %s" % p)
    cols = read_csv(op_path)
    for c, name in enumerate(component_types(config)):
        re_key, im_key = name + "_re", name + "_im"
        if re_key not in cols or im_key not in cols:
            raise KeyError("the file does not contain the columns '%s' and '%s'"
                           % (re_key, im_key))
        re = _square(cols[re_key], n, "order parameter")
        im = _square(cols[im_key], n, "order parameter")
        sim.set_order_parameter(c, re + 1j * im)
    cols = read_csv(a_path)
    if "a_x" not in cols or "a_y" not in cols:
        raise KeyError("the file does not contain the columns 'a_x' and 'a_y'")
    ax = _square(cols["a_x"], n, "vector potential")
    ay = _square(cols["a_y"], n, "vector potential")
    if config["physics"].get("penetration_depth", -1.0) > 0:
        sim.set_vector_potential(ax, ay)
    else:
        sim.set_vector_potential(np.zeros_like(ax), np.zeros_like(ay))


# ----------------------------------------------------------------- writing
def write_grid_csv(path, titles, planes, integer=False):
    """CSVGridWriter::write: `planes` is a list with one entry per title,
    each a real (n, n) array or a complex one (written as _re, _im)."""
    n = planes[0].shape[0]
    header = ["x_idx", "y_idx"]
    cols = []
    for t, p in zip(titles, planes):
        if np.iscomplexobj(p):
            header += [t + "_re", t + "_im"]
            cols += [np.ascontiguousarray(p.real), np.ascontiguousarray(p.imag)]
        else:
            header.append(t)
            cols.append(np.asarray(p))
    yy, xx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    with open(path, "w") as f:
        f.write(",".join(header) + "\n")
        flat = [c.reshape(-1) for c in cols]
        xs, ys = xx.reshape(-1), yy.reshape(-1)
        lines = []
        for i in range(n * n):
            vals = ["%d" % int(c[i]) if integer else _fmt(c[i]) for c in flat]
            lines.append("%d,%d,%s" % (xs[i], ys[i], ",".join(vals)))
        f.write("\n".join(lines) + "\n")


def curl(ax, ay, h):
    """GridFunctions curl_kernel (GridFunctions_GPU.h:327-356): one-sided at
    the edges, central inside."""
    n = ax.shape[0]
    idx = np.arange(n)
    nxt = np.minimum(idx + 1, n - 1)
    prv = np.maximum(idx - 1, 0)
    den = (nxt - prv).astype(float)
    day_dx = (ay[:, nxt] - ay[:, prv]) / den[None, :]
    dax_dy = (ax[nxt, :] - ax[prv, :]) / den[:, None]
    return (day_dx - dax_dy) / h


def moment(jx, jy, h):
    """GridFunctions::moment (GridFunctions_CPU.h:291-320): 0.5 h^3 sum r x j."""
    n = jx.shape[0]
    c = 0.5 * (n - 1)
    y, x = np.meshgrid(np.arange(n) - c, np.arange(n) - c, indexing="ij")
    return 0.5 * h ** 3 * float(np.sum(x * jy - y * jx))


def save_state(sim, config, save_path):
    """SimulationWriter::write (io/SimulationWriter.h:50-236), CSV format."""
    os.makedirs(save_path, exist_ok=True)
    d = sim.dims
    n, h = sim.N0, d.h
    res = sim.results()
    domain = sim.domain()
    delta = sim.order_parameter(1)       # ComputeOrderParameter::getResult
    j = sim.current()
    a = sim.vector_potential(0)          # unscaled induced potential
    lam = config["physics"].get("penetration_depth", -1.0)
    if not lam > 0:
        a = np.zeros_like(a)
    bz = curl(a[0], a[1], h)
    write_grid_csv(os.path.join(save_path, "domain.csv"), ["domain"], [domain],
                   integer=True)
    write_grid_csv(os.path.join(save_path, "order_parameter.csv"),
                   component_types(config), list(delta))
    write_grid_csv(os.path.join(save_path, "current_density.csv"), ["j_x", "j_y"],
                   [j[0], j[1]])
    write_grid_csv(os.path.join(save_path, "vector_potential.csv"), ["a_x", "a_y"],
                   [a[0], a[1]])
    write_grid_csv(os.path.join(save_path, "magnetic_flux_density.csv"), ["b_z"], [bz])
    ng = d.num_grain_sites
    area = h * h * ng
    inside = domain.astype(bool)
    results = {
        "free_energy_per_unit_area": res["free_energy_per_area"],
        "magnetic_free_energy_per_unit_area": res["magnetic_free_energy_per_area"],
        "order_parameter_average_magnitude": float(np.sum(np.abs(delta))) / ng,
        "current_density_average_magnitude": float(np.sum(np.hypot(j[0], j[1]))) / ng,
        "magnetic_moment_per_unit_area": moment(j[0], j[1], h) / area,
        "flux_density_average": float(np.sum(np.where(inside, bz, 0.0))) / ng,
        "vector_potential_average_magnitude":
            float(np.sum(np.where(inside, np.hypot(a[0], a[1]), 0.0))) / ng,
    }
    with open(os.path.join(save_path, "simulation_results.json"), "w") as f:
        json.dump(results, f, indent=4)
    internal = {"area_compute": area, "internal_lattice_points": ng,
                "grain_margin": GRAIN_SIZE_MARGIN, "num_lattice_points": n * n}
    with open(os.path.join(save_path, "internal_parameters.json"), "w") as f:
        json.dump(internal, f, indent=4)
    return results


class ConvergenceWriter:
    """CSVSimulationIterationWriter: one row per iteration, appended."""

    HEADER = ("iteration,boundary_coherence_residual,order_parameter_residual,"
              "current_density_residual,vector_potential_residual,"
              "free_energy_residual,free_energy_per_unit_area")

    def __init__(self, path):
        self.path = path

    def write(self, res):
        new = not os.path.exists(self.path)
        with open(self.path, "a") as f:
            if new:
                f.write(self.HEADER + "\n")
            f.write("%d,%s\n" % (res["iterations"], ",".join(_fmt(res[k]) for k in (
                "residual_boundary", "residual_order_parameter", "residual_current",
                "residual_vector_potential", "residual_free_energy",
                "free_energy_per_area"))))


def _check_data_format(misc):
    fmt = misc.get("data_format", "csv")
    if fmt not in ("csv", "h5"):
        raise ValueError("misc.data_format must be 'csv' or 'h5', got %r" % (fmt,))
    if fmt == "h5" and misc.get("save_path", "") and misc.get("save_frequency", -1) != 0:
        raise ValueError("misc.data_format 'h5' (HDF5 output) is not supported by this "
                         "build: set misc.data_format to 'csv' (the reference writes the "
                         "same grids as CSV, io/CSVGridWriter.h)")


def _run_iterations(step, sim, compute, n_min, n_max, criterion, writer, verbose):
    """runIterations (run_iteration.h:85-141): iterate until converged with
    the minimum reached, or to the maximum (< 0: none); burn-in iterations
    converge on the boundary residual.  Every iteration writes a row."""
    if n_min == 0 and n_max == 0:
        return False
    i = 1
    while True:
        step()
        r = sim.results()
        conv = bool(r["converged"]) if compute else r["residual_boundary"] < criterion
        if verbose:
            print("-- iteration %d residual %.3e free energy %.10f" % (
                r["iterations"], r["residual"], r["free_energy_per_area"]))
        if writer:
            writer.write(r)
        if (conv and i >= n_min) or (n_max >= 0 and i >= n_max):
            return conv
        i += 1


def simulate(config, precision="fp64", verbose=False):
    """The reference's simulation binary in process (src/simulation.cu,
    runSimulation of run_simulation.h:124-187): read the config, warm start
    from misc.load_path, burn in, iterate with the min/max limits, write
    misc.save_path.  Returns (converged, results dict)."""
    import superconga_amd as sa
    if isinstance(config, str):
        with open(config) as f:
            config = json.load(f)
    misc = config.get("misc", {})
    _check_data_format(misc)
    nu = config["numerics"]
    criterion = nu.get("convergence_criterion", 1e-6)
    sim = sa.Simulation(config)
    try:
        load_path = misc.get("load_path", "")
        if load_path:
            load_state(sim, config, load_path)
        save_path = misc.get("save_path", "")
        save_freq = misc.get("save_frequency", -1)
        save = bool(save_path) and save_freq != 0
        writer = None
        if save:
            os.makedirs(save_path, exist_ok=True)
            with open(os.path.join(save_path, "simulation_config.json"), "w") as f:
                json.dump(config, f, indent=4)
            writer = ConvergenceWriter(os.path.join(save_path, "simulation_convergence.csv"))
        n_burn = nu.get("num_iterations_burnin", 0)
        n_min = nu.get("num_iterations_min", 0)
        n_max = nu.get("num_iterations_max", -1)
        _run_iterations(lambda: sim.burnin(precision), sim, False, n_burn, n_burn, criterion,
                        writer, verbose)
        converged = False
        if save_freq > 0:
            # one iteration at a time; the limits and the save schedule count
            # the burn-in iterations too (Context::getNumIterations)
            while True:
                converged = _run_iterations(lambda: sim.compute(precision), sim, True, 0, 1,
                                            criterion, writer, verbose)
                it = sim.results()["iterations"]
                if (converged and it >= n_min) or (n_max >= 0 and it >= n_max):
                    break
                if save and it % save_freq == 0:
                    save_state(sim, config, save_path)
        else:
            converged = _run_iterations(lambda: sim.compute(precision), sim, True, n_min, n_max,
                                        criterion, writer, verbose)
        results = save_state(sim, config, save_path) if save else None
        return converged, results
    finally:
        sim.close()


def write_ldos_csv(path, ldos):
    """CSVGridWriter::write of the LDOS grid with the single title "ldos":
    the energy index becomes a column (io/CSVGridWriter.h, fieldsAsIndex)."""
    ne, n, _ = ldos.shape
    yy, xx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    xs, ys = xx.reshape(-1), yy.reshape(-1)
    with open(path, "w") as f:
        f.write("x_idx,y_idx,e_idx,ldos\n")
        for e in range(ne):
            vals = ldos[e].reshape(-1)
            f.write("\n".join("%d,%d,%d,%s" % (xs[i], ys[i], e, _fmt(vals[i]))
                              for i in range(n * n)) + "\n")


class PostprocessConvergenceWriter:
    """CSVPostprocessIterationWriter: iteration, boundary and LDOS residuals."""

    HEADER = "iteration,boundary_coherence_residual,ldos_residual"

    def __init__(self, path):
        self.path = path

    def write(self, iteration, st):
        new = not os.path.exists(self.path)
        with open(self.path, "a") as f:
            if new:
                f.write(self.HEADER + "\n")
            f.write("%d,%s,%s\n" % (iteration, _fmt(st["residual_boundary"]),
                                     _fmt(st["residual_ldos"])))


def postprocess(config, precision="fp64", verbose=False):
    """The reference's postprocess binary in process: read the postprocess
    config and misc.load_path/simulation_config.json, load the order
    parameter and vector potential written there, run the Keldysh iteration
    block by block and write misc.save_path.  Returns (converged, ldos)."""
    import superconga_amd as sa
    if isinstance(config, str):
        with open(config) as f:
            config = json.load(f)
    misc = config.get("misc", {})
    _check_data_format(misc)
    load_path = misc.get("load_path", "")
    sim_path = os.path.join(load_path, "simulation_config.json")
    if not os.path.exists(sim_path):
        raise FileNotFoundError("-- ERROR! Could not find/open config file '%s'." % sim_path)
    with open(sim_path) as f:
        sim_cfg = json.load(f)
    nu = config["numerics"]
    spec = sa.Spectroscopy(sim_cfg, config)
    try:
        load_state(spec, sim_cfg, load_path)
        save_path = misc.get("save_path", "")
        writer = None
        if save_path:
            os.makedirs(save_path, exist_ok=True)
            conv_path = os.path.join(save_path, "postprocess_convergence.csv")
            if os.path.exists(conv_path):
                os.remove(conv_path)
            writer = PostprocessConvergenceWriter(conv_path)

        def iterations(compute, mn, mx):
            # runIterations (run_iteration.h:85-141)
            if mn == 0 and mx == 0:
                return False
            conv = False
            i = 1
            while True:
                spec.compute(precision, burnin=not compute)
                st = spec.status()
                conv = bool(st["partially_converged"]) if compute else (
                    st["residual_boundary"] < nu["convergence_criterion"])
                if writer:
                    writer.write(st["iterations"], st)
                if verbose:
                    print("Iteration: %d | Residual: %e" % (st["iterations"], st["residual"]))
                if (conv and i >= mn) or (mx >= 0 and i >= mx):
                    break
                i += 1
            return conv

        # runPostprocess (run_simulation.h:44-108)
        nb = nu.get("num_iterations_burnin", -1)
        converged = False
        while not converged:
            iterations(False, nb, nb)
            block_converged = iterations(True, nu.get("num_iterations_min", 0),
                                         nu.get("num_iterations_max", -1))
            converged = bool(spec.status()["converged"])
            if not block_converged:
                break
        ldos = spec.ldos()
        if save_path:
            with open(os.path.join(save_path, "postprocess_config.json"), "w") as f:
                json.dump(config, f, indent=4)
            write_ldos_csv(os.path.join(save_path, "ldos.csv"), ldos)
        return converged, ldos
    finally:
        spec.close()
