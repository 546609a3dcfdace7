"""The reference's on-disk formats (superconga_amd.io): CSV grids as
io/CSVGridWriter.h writes them and io/CSVReader.h reads them, the
convergence log of io/CSVSimulationIterationWriter.h, and the derived
quantities of io/SimulationWriter.h."""
import os

import numpy as np
import pytest

from superconga_amd import io as scio


def test_grid_csv_layout_and_roundtrip(tmp_path):
    n = 5
    rng = np.random.default_rng(0)
    re = rng.normal(size=(n, n))
    cx = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    p = tmp_path / "g.csv"
    scio.write_grid_csv(str(p), ["a_x", "s"], [re, cx])
    lines = p.read_text().splitlines()
    assert lines[0] == "x_idx,y_idx,a_x,s_re,s_im"
    assert len(lines) == 1 + n * n
    # row-major, x fastest; std::scientific with 8 digits
    assert lines[1].startswith("0,0,")
    assert lines[2].startswith("1,0,")
    assert lines[1 + n].startswith("0,1,")
    assert lines[1].split(",")[2] == "%.8e" % re[0, 0]
    back = scio.read_csv(str(p))
    assert set(back) == {"a_x", "s_re", "s_im"}
    assert np.allclose(back["a_x"].reshape(n, n), re, rtol=1e-8, atol=0)
    assert np.allclose((back["s_re"] + 1j * back["s_im"]).reshape(n, n), cx, rtol=1e-8,
                       atol=0)


def test_integer_grid(tmp_path):
    d = np.zeros((4, 4), dtype=np.int8)
    d[1:3, 1:3] = 1
    p = tmp_path / "domain.csv"
    scio.write_grid_csv(str(p), ["domain"], [d], integer=True)
    lines = p.read_text().splitlines()
    assert lines[0] == "x_idx,y_idx,domain"
    assert lines[1 + 4 + 1] == "1,1,1"
    assert np.array_equal(scio.read_csv(str(p))["domain"].reshape(4, 4), d)


def test_curl_of_symmetric_gauge_is_uniform():
    # A = B/2 (-y, x): curl A = B everywhere, also at the one-sided edges.
    n, h, b = 9, 0.2, 1.7
    c = 0.5 * (n - 1)
    y, x = np.meshgrid((np.arange(n) - c) * h, (np.arange(n) - c) * h, indexing="ij")
    bz = scio.curl(-0.5 * b * y, 0.5 * b * x, h)
    assert np.allclose(bz, b, rtol=1e-13)


def test_moment_of_rigid_rotation():
    # j = (-y, x) on an n x n grid: moment = 0.5 h^3 sum (x^2 + y^2)
    n, h = 7, 0.5
    c = 0.5 * (n - 1)
    y, x = np.meshgrid(np.arange(n) - c, np.arange(n) - c, indexing="ij")
    m = scio.moment(-y, x, h)
    assert abs(m - 0.5 * h ** 3 * np.sum(x * x + y * y)) < 1e-12


def test_convergence_writer(tmp_path):
    p = tmp_path / "simulation_convergence.csv"
    w = scio.ConvergenceWriter(str(p))
    r = {"iterations": 3, "residual_boundary": 1e-3, "residual_order_parameter": 2e-3,
         "residual_current": 3e-3, "residual_vector_potential": 4e-3,
         "residual_free_energy": 5e-3, "free_energy_per_area": -0.0185}
    w.write(r)
    w.write(dict(r, iterations=4))
    lines = p.read_text().splitlines()
    assert lines[0] == scio.ConvergenceWriter.HEADER
    assert lines[1] == "3,1.00000000e-03,2.00000000e-03,3.00000000e-03,4.00000000e-03," \
                       "5.00000000e-03,-1.85000000e-02"
    assert lines[2].startswith("4,")


def test_component_types_follow_reference_order():
    cfg = {"physics": {"order_parameter": {"s": {}, "dx2-y2": {}}}}
    assert scio.component_types(cfg) == ["dx2-y2", "s"]


def test_load_rejects_other_lattice(tmp_path):
    class Fake:
        N0 = 4
    d = np.ones((3, 3))
    scio.write_grid_csv(str(tmp_path / "domain.csv"), ["domain"], [d.astype(np.int8)],
                        integer=True)
    scio.write_grid_csv(str(tmp_path / "order_parameter.csv"), ["s"], [d + 0j])
    scio.write_grid_csv(str(tmp_path / "vector_potential.csv"), ["a_x", "a_y"], [d, d])
    cfg = {"physics": {"order_parameter": {"s": {}}}}
    with pytest.raises(NotImplementedError):
        scio.load_state(Fake(), cfg, str(tmp_path))


def test_ldos_csv_energy_index_layout(tmp_path):
    # CSVGridWriter with one title for a multi-field grid: e_idx column,
    # fields outermost, then rows, x fastest (io/CSVGridWriter.h)
    ne, n = 3, 4
    ldos = np.arange(ne * n * n, dtype=float).reshape(ne, n, n) / 7.0
    p = tmp_path / "ldos.csv"
    scio.write_ldos_csv(str(p), ldos)
    lines = p.read_text().splitlines()
    assert lines[0] == "x_idx,y_idx,e_idx,ldos"
    assert len(lines) == 1 + ne * n * n
    assert lines[1] == "0,0,0,%.8e" % ldos[0, 0, 0]
    assert lines[2].startswith("1,0,0,")
    assert lines[1 + n].startswith("0,1,0,")
    assert lines[1 + n * n] == "0,0,1,%.8e" % ldos[1, 0, 0]
    back = scio.read_csv(str(p))
    assert np.allclose(back["ldos"].reshape(ne, n, n), ldos, rtol=1e-8, atol=0)


def test_postprocess_convergence_writer(tmp_path):
    p = tmp_path / "postprocess_convergence.csv"
    w = scio.PostprocessConvergenceWriter(str(p))
    w.write(5, {"residual_boundary": 1e-3, "residual_ldos": 2e-4})
    lines = p.read_text().splitlines()
    assert lines == ["iteration,boundary_coherence_residual,ldos_residual",
                     "5,1.00000000e-03,2.00000000e-04"]


def test_hdf5_output_is_refused_before_any_device_work(tmp_path):
    cfg = {"physics": {}, "numerics": {},
           "misc": {"data_format": "h5", "save_path": str(tmp_path), "save_frequency": -1}}
    with pytest.raises(ValueError, match="h5"):
        scio.simulate(cfg)
    with pytest.raises(ValueError, match="data_format"):
        scio.simulate({"physics": {}, "numerics": {}, "misc": {"data_format": "xml"}})


class _FakeSim:
    """Residual sequences standing in for a grain: burn-in iterations
    converge on the boundary residual, compute iterations on `converged`."""

    def __init__(self, boundary, converged):
        self.boundary = list(boundary)
        self.conv = list(converged)
        self.it = 0
        self.last = None

    def burnin(self):
        self.it += 1
        self.last = {"iterations": self.it, "residual": 1.0, "free_energy_per_area": 0.0,
                     "residual_boundary": self.boundary.pop(0), "converged": 0}

    def compute(self):
        self.it += 1
        self.last = {"iterations": self.it, "residual": 1.0, "free_energy_per_area": 0.0,
                     "residual_boundary": 0.0, "converged": self.conv.pop(0)}

    def results(self):
        return self.last


class _Rows:
    def __init__(self):
        self.rows = []

    def write(self, r):
        self.rows.append(r["iterations"])


def test_run_iterations_limits_like_the_reference():
    # runIterations (run_iteration.h:85-141): burn-in of n runs exactly n
    # iterations, -1 runs until the boundary residual is below the criterion,
    # and every iteration (burn-in included) writes a row
    sim, rows = _FakeSim([1.0, 0.5, 1e-9, 1.0], []), _Rows()
    assert scio._run_iterations(sim.burnin, sim, False, -1, -1, 1e-6, rows, False)
    assert rows.rows == [1, 2, 3]
    sim, rows = _FakeSim([1e-9, 1e-9, 1e-9], []), _Rows()
    scio._run_iterations(sim.burnin, sim, False, 2, 2, 1e-6, rows, False)
    assert rows.rows == [1, 2]
    assert not scio._run_iterations(sim.burnin, sim, False, 0, 0, 1e-6, rows, False)
    # compute: converged but the minimum not reached keeps going
    sim, rows = _FakeSim([], [1, 1, 1, 0]), _Rows()
    assert scio._run_iterations(sim.compute, sim, True, 3, -1, 1e-6, rows, False)
    assert rows.rows == [1, 2, 3]
    sim, rows = _FakeSim([], [0] * 5), _Rows()
    assert not scio._run_iterations(sim.compute, sim, True, 0, 4, 1e-6, rows, False)
    assert rows.rows == [1, 2, 3, 4]
