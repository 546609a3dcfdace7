"""The C ABI without a device: the library loads, exports every entry point
include/superconga_amd.h declares, reports errors the documented way, and
its host-side planning (lattice sizes, Ozaki tables, geometry labels,
boundary-storage sizes) agrees exactly with the oracle."""
import ctypes
import json
import math
import os
import re

import numpy as np
import pytest

import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "superconga_amd.h")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sc_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol(product):
    names = declared_symbols()
    assert len(names) >= 25
    L = ctypes.CDLL(product.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing


def test_version_and_error_reporting(product):
    L = product.lib()
    assert L.sc_version() >= 1
    h = ctypes.c_void_p()
    code = L.sc_create(b"{not json", ctypes.byref(h))
    assert code != 0
    assert L.sc_last_error()
    code = L.sc_create(json.dumps({"physics": {}}).encode(), ctypes.byref(h))
    assert code != 0
    assert L.sc_get_dims(None, None) != 0
    assert L.sc_ozaki(0, None, None) != 0


def test_describe_rejects_bad_angle(product):
    with pytest.raises(product.SupercongaError):
        product.describe(cases.base(angles=8), angle=8)


def test_host_tables_match_oracle(product, oracle):
    table = json.load(open(os.path.join(GOLDEN, "ozaki_max_poles.json")))["max_pole"]
    for n in (1, 2, 3, 9, 14, 28, 61, 200):
        p, r = product.ozaki(n)
        po, ro = oracle.ozaki(n)
        assert abs(p.max() - table[n - 1]) < 1e-9 * table[n - 1]
        # Different eigen-solvers (QL here, Jacobi in the oracle): agree to
        # rounding of the eigenproblem.
        # (3.9e-12 relative on the residues at 200 poles, 1.2e-13 at 28).
        assert np.allclose(np.sort(p), np.sort(po), rtol=1e-11, atol=0)
        assert np.allclose(r[np.argsort(p)], ro[np.argsort(po)], rtol=2e-11, atol=0)
        # Ozaki's expansion of the Fermi function reproduces tanh(x/2).
        if n >= 9:
            for x in (0.1, 1.0, 5.0, 15.0):
                assert abs(np.sum(4 * x * r / (x * x + p * p)) - math.tanh(x / 2)) < 1e-12
    for t, cut in ((0.5, 10 / (2 * math.pi)), (0.5, 16.0), (0.2, 16.0), (0.1, 32.0),
                   (0.3, 16.0), (0.9, 16.0), (0.01, 64.0)):
        assert product.num_ozaki(t, cut) == oracle.num_ozaki(t, cut)
    for t in (0.001, 0.1, 0.5, 0.9, 0.999, 1.0):
        assert abs(product.bulk_gap(t) - oracle.bulk_gap(t)) < 1e-13


PLAN_CASES = {
    "disc": cases.base(angles=16),
    "square_rotated_axes": cases.base(
        order_parameter={"dx2-y2": {"critical_temperature": 1.0}}, rotation=0.125,
        components=[cases.regular_polygon_unit(4)], angles=20, ppxi=4.0),
    "octagon_hole": cases.mixed_small(),
    "polygon": cases.base(components=[{"polygon": {
        "vertices_x": [-0.5, 0.5, 0.3, -0.4], "vertices_y": [-0.45, -0.5, 0.5, 0.2]}}],
        angles=12, ppxi=4.0, width=12.0),
    "disc_7_angles": cases.base(angles=7, ppxi=3.0),
    "square_18_angles": cases.base(angles=18, ppxi=3.0,
                                   components=[cases.regular_polygon_unit(4)]),
    "two_discs": cases.base(components=[
        {"disc": {"radius": 0.25, "center_x": -0.22, "center_y": 0.0}},
        {"disc": {"radius": 0.25, "center_x": 0.22, "center_y": 0.1}}],
        angles=16, ppxi=3.0, width=14.0),
    "physical_units_square": cases.base(geometry=[
        {"regular_polygon": {"num_edges": 4, "side_length": 8.0, "rotation": 0.1}}],
        angles=12, ppxi=3.0),
}


@pytest.mark.parametrize("name", sorted(PLAN_CASES))
def test_plan_matches_oracle(product, oracle, name):
    cfg = PLAN_CASES[name]
    d, _ = product.describe(cfg)
    o = oracle.Oracle(cfg)
    assert (d.n0, d.n1, d.num_energies, d.num_angles, d.num_components) == (
        o.N0, o.N1, o.Ne, o.Na, o.Nc)
    assert d.num_grain_sites == o.get_int("num_grain_sites")
    assert d.num_in_gamma == o.get_int("num_storage_in_gamma")
    assert d.num_out_gamma == o.get_int("num_storage_out_gamma")
    assert d.num_in_gammatilde == o.get_int("num_storage_in_gammatilde")
    assert d.num_out_gammatilde == o.get_int("num_storage_out_gammatilde")
    assert d.traj_steps == o.get_int("traj_steps")
    for k in range(o.Na):
        _, lab = product.describe(cfg, angle=k)
        assert np.array_equal(lab, o.geometry(k)[0]), k


def test_plan_octagon_benchmark_sizes(product):
    d, _ = product.describe(cases.octagon_dwave())
    assert (d.n0, d.n1, d.num_energies, d.num_angles) == (930, 1316, 28, 128)
    # About 2 Na Ne N_grain steps: the in-grain share of the frames.
    approx = 2 * d.num_angles * d.num_energies * d.num_grain_sites
    assert 0.9 < d.traj_steps / approx < 1.2


EXAMPLES = os.path.join(GOLDEN, "examples")


def example_config(name):
    with open(os.path.join(EXAMPLES, name + ".json")) as f:
        cfg = json.load(f)
    # unchanged: every example asks for the eigen-truncated Poisson kernel
    # (vector_potential_error 1e-6), which the product and the oracle build
    return cfg


@pytest.mark.parametrize("name", sorted(n[:-5] for n in os.listdir(EXAMPLES)))
def test_reference_example_configs_plan_like_oracle(product, oracle, name):
    """The reference's own example configs (examples/*/simulation_config.json)
    are accepted unchanged and planned exactly like the oracle plans them."""
    cfg = example_config(name)
    d, _ = product.describe(cfg)
    o = oracle.Oracle(cfg)
    assert (d.n0, d.n1, d.num_energies, d.num_angles, d.num_components) == (
        o.N0, o.N1, o.Ne, o.Na, o.Nc)
    assert d.num_grain_sites == o.get_int("num_grain_sites")
    assert d.traj_steps == o.get_int("traj_steps")
    assert (d.num_in_gamma, d.num_out_gamma, d.num_in_gammatilde, d.num_out_gammatilde) == (
        o.get_int("num_storage_in_gamma"), o.get_int("num_storage_out_gamma"),
        o.get_int("num_storage_in_gammatilde"), o.get_int("num_storage_out_gammatilde"))
    for k in sorted({0, o.Na // 8, o.Na // 3, o.Na - 1}):
        _, lab = product.describe(cfg, angle=k)
        assert np.array_equal(lab, o.geometry(k)[0]), k


def test_spectroscopy_config_errors(product):
    # parseSpectroscopy rejects a postprocess config before touching a device
    L = product.lib()
    sim = json.dumps(cases.base(angles=8)).encode()
    good = {"spectroscopy": {"energy_min": 0.0, "energy_max": 0.5, "num_energies": 8,
                             "energy_broadening": 0.01},
            "numerics": {"convergence_criterion": 1e-4, "norm": "l2",
                         "num_energies_per_block": 4, "num_fermi_momenta": 8}}
    h = ctypes.c_void_p()
    for key, sub, bad in (("spectroscopy", "num_energies", 0),
                          ("spectroscopy", "energy_broadening", 0.0),
                          ("numerics", "num_energies_per_block", 0),
                          ("numerics", "num_fermi_momenta", 0),
                          ("numerics", "norm", "l3")):
        cfg = json.loads(json.dumps(good))
        cfg[key][sub] = bad
        code = L.sc_ldos_create(sim, json.dumps(cfg).encode(), ctypes.byref(h))
        assert code == 2, (sub, code)  # SC_ERR_CONFIG
        assert L.sc_last_error()
    assert L.sc_ldos_create(sim, b"{broken", ctypes.byref(h)) != 0
    assert L.sc_ldos_compute(None, 0, 0) == 1      # SC_ERR_ARGUMENT: null context
    assert L.sc_newton_krylov(None, 1, 4, ctypes.c_double(1e-8), None) == 1
    # the entry points of this build without reference counterpart refuse a
    # null handle before any device work
    assert L.sc_newton_krylov_ex(None, 1, 4, ctypes.c_double(1e-8), 1, None) == 1
    assert L.sc_batch_newton_krylov(None, 1, 4, ctypes.c_double(1e-8), 1, None, None) == 1
    assert L.sc_set_graphs(None, 1) == 1
    assert L.sc_batch_set_graphs(None, 1) == 1
    assert L.sc_set_screening(None, ctypes.c_double(0.5)) == 1
    assert L.sc_free_energy_gradient_ex(None, None, None, None) == 1
    assert L.sc_residual_gradient(None, None, None, None) == 1
    assert L.sc_jacobian_vector(None, None, None) == 1
    assert L.sc_set_adjoint_checkpoints(None, 0) == 1
    assert L.sc_set_newton_options(None, 4, ctypes.c_double(0.3), 1, 0) == 1


def test_noalloc_poisson_refuses_oversized_grid(product):
    # refused before any device work: the static scratch covers n <= 512
    import ctypes
    L = product.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    L.sc_poisson_solve_noalloc.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                           ctypes.c_int, dp, dp]
    j = np.zeros(2 * 600 * 600)
    out = np.zeros(2 * 600 * 600)
    rc = L.sc_poisson_solve_noalloc(600, 0.05, 100, 1, j.ctypes.data_as(dp),
                                    out.ctypes.data_as(dp))
    assert rc == 1
    assert "600" in L.sc_last_error().decode()
