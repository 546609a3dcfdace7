"""Keldysh spectroscopy on the GPU (sc_ldos_*) against the oracle's
restatement of src/postprocess.cu, to 1e-10 of the LDOS maximum in FP64.

The oracle's LDOS is pinned by the closed forms of test_oracle_ldos.py
(broadened BCS density of states for a uniform s-wave gap, one in the normal
state, no coupling between energy blocks); the same BCS check runs here on
the product directly.
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

TOL = 1e-10
SPEC = {"energy_min": 0.0, "energy_max": 0.5, "num_energies": 6, "energy_broadening": 0.02}


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def postprocess_config(per_block=3, criterion=1e-9, angles=16, spec=SPEC):
    return {"spectroscopy": dict(spec),
            "numerics": {"convergence_criterion": criterion, "norm": "l2",
                         "num_energies_per_block": per_block, "num_fermi_momenta": angles,
                         "num_iterations_burnin": 0, "num_iterations_max": 300,
                         "num_iterations_min": 0},
            "misc": {"load_path": "", "save_path": ""}}


def converged_state(product, cfg, iterations=6):
    """An order parameter (and vector potential) a few iterations into a
    simulation: the state the LDOS is computed for."""
    s = product.Simulation(cfg)
    for _ in range(iterations):
        s.compute("fp64")
    delta = s.order_parameter(0)
    a = s.vector_potential(0)
    s.close()
    return delta, a


def ldos_pair(product, oracle, cfg, post, delta, a):
    sp = product.Spectroscopy(cfg, post)
    o = oracle.Oracle(oracle.postprocess_config(cfg, post))
    for c, d in enumerate(delta):
        sp.set_order_parameter(c, d)
        o.set_order_parameter(c, d)
    # both take the unscaled induced potential
    sp.set_vector_potential(a[0], a[1])
    o.set_vector_potential(a[0], a[1])
    nu = post["numerics"]
    o.ldos_setup(post["spectroscopy"], nu["num_energies_per_block"])
    return sp, o


CASES = {
    # s-wave vortex in a disc with field and induced vector potential
    "vortex_field": (lambda: cases.base(
        temperature=0.4, flux=1.5, lam=4.0, angles=16, ppxi=3.0, criterion=1e-9,
        order_parameter={"s": {"critical_temperature": 1.0,
                               "vortices": [{"center_x": 0.0, "center_y": 0.0,
                                             "winding_number": 1}]}})),
    # d-wave hexagon with 45-degree crystal axes: pair-breaking edges
    "dwave_hexagon": (lambda: cases.base(
        temperature=0.3, angles=16, ppxi=3.0, criterion=1e-9,
        order_parameter={"dx2-y2": {"critical_temperature": 1.0}}, rotation=0.125,
        components=[cases.regular_polygon_unit(6)])),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_ldos_matches_oracle(product, oracle, name):
    cfg = CASES[name]()
    cfg["numerics"]["norm"] = "l2"
    delta, a = converged_state(product, cfg)
    post = postprocess_config(per_block=3)
    sp, o = ldos_pair(product, oracle, cfg, post, delta, a)
    sp.set_angle_chunk(5)  # several chunks of spectral weights
    # iteration by iteration on the first block
    for it in range(3):
        sp.compute("fp64", burnin=(it == 0))
        o.ldos_compute(burnin=(it == 0))
        st = sp.status()
        err_b = abs(st["residual_boundary"] - o.get("residual_boundary")) / o.get(
            "residual_boundary")
        print(name, "it", it, "boundary residual %.3e err %.1e" % (st["residual_boundary"], err_b))
        assert err_b < TOL
        if it > 0:
            assert abs(st["residual_ldos"] - o.get("residual_ldos")) <= TOL * o.get("residual_ldos")
    e = rel(sp.ldos(), o.ldos())
    print(name, "LDOS after 3 iterations: rel err %.2e" % e)
    assert e < TOL
    # then the whole runPostprocess, block by block
    conv = sp.run(0, 0, 300, "fp64")
    oconv = o.ldos_run(0, 0, 300)
    st = sp.status()
    assert conv and oconv
    assert st["iterations"] == o.get_int("num_iterations")
    assert st["converged"]
    e = rel(sp.ldos(), o.ldos())
    print(name, "converged LDOS: %d iterations, rel err %.2e" % (st["iterations"], e))
    assert e < TOL
    sp.close()


def test_ldos_bcs_closed_form(product):
    cfg = cases.base(temperature=0.5, angles=16, ppxi=3.0, criterion=1e-10)
    cfg["numerics"]["norm"] = "l2"
    spec = dict(SPEC, num_energies=9, energy_broadening=0.01)
    post = postprocess_config(per_block=4, criterion=1e-10, spec=spec)
    sp = product.Spectroscopy(cfg, post)
    d0 = 0.2
    sp.set_order_parameter(0, np.full((sp.N0, sp.N0), d0, dtype=complex))
    assert sp.run(-1, 0, 50, "fp64")
    ldos = sp.ldos()
    inside = sp.domain() > 0
    eps = sp.energies_ldos() + 1j * spec["energy_broadening"]
    bcs = (-1j * eps / np.sqrt(d0 ** 2 - eps ** 2)).real
    for e in range(spec["num_energies"]):
        assert np.max(np.abs(ldos[e][inside] - bcs[e])) < 1e-11 * max(bcs[e], 1.0), e
        assert np.all(ldos[e][~inside] == 0.0)
    sp.close()


@pytest.mark.parametrize("symmetry", ["s", "dx2-y2"])
def test_reference_dos_known_answer(product, symmetry):
    # tests/physics/dos_test.cu:32-241 on the product (see test_oracle_ldos.py)
    from test_oracle_ldos import boole_weight
    n_angles = 160
    comps = [{"disc": {"radius": 0.5}}] if symmetry == "s" else [
        {"regular_polygon": {"num_edges": 4}}]
    cfg = cases.base(temperature=0.5, angles=n_angles, ppxi=2.0, criterion=1e-8,
                     norm="linf", width=5.0, components=comps,
                     order_parameter={symmetry: {"critical_temperature": 1.0}})
    spec = {"energy_min": 0.0, "energy_max": 0.5, "num_energies": 32,
            "energy_broadening": 0.008}
    post = postprocess_config(per_block=-1, criterion=1e-8, angles=n_angles, spec=spec)
    post["numerics"]["norm"] = "linf"
    sim = product.Simulation(cfg)  # its initial order parameter is the bulk one
    delta = sim.order_parameter(0)[0]
    sim.close()
    sp = product.Spectroscopy(cfg, post)
    sp.set_order_parameter(0, delta)
    assert sp.run(0, 0, 42, "fp64")
    ldos = sp.ldos()
    inside = sp.domain() > 0
    d0 = float(np.max(np.abs(delta)))
    eps = np.linspace(0.0, 0.5, 32) + 0.008j
    phi = np.pi * 2 * np.arange(n_angles) / n_angles + 0.5 * np.pi
    w = np.array([boole_weight(k, n_angles) for k in range(n_angles)]) / n_angles
    if symmetry == "s":
        eta = np.ones(n_angles)
    else:
        raw = np.cos(phi) ** 2 - np.sin(phi) ** 2
        eta = raw / np.sqrt(np.sum(w * raw * raw))
    for e in range(32):
        omega = np.sqrt(np.abs(d0 * eta) ** 2 - eps[e] ** 2)
        expected = float(np.sum(w * (eps[e] / omega).imag))
        got = float(np.sum(ldos[e][inside]) / np.count_nonzero(inside))
        assert abs(got - expected) <= 1e-8 * abs(expected), (e, got, expected)
    sp.close()


def test_ldos_chunking_and_shapes_agree(product):
    cfg = CASES["dwave_hexagon"]()
    cfg["numerics"]["norm"] = "l2"
    delta, a = converged_state(product, cfg)
    post = postprocess_config(per_block=-1)
    out = []
    for chunk, shape in ((0, None), (1, None), (7, (4, 2, 8)), (16, (16, 1, 4))):
        sp = product.Spectroscopy(cfg, post)
        sp.set_order_parameter(0, delta[0])
        sp.set_angle_chunk(chunk)
        if shape:
            sp.set_sweep_shape(*shape, precision="fp64")
        for it in range(4):
            sp.compute("fp64")
        out.append(sp.ldos())
        sp.close()
    for other in out[1:]:
        # the same arithmetic in a different order of launches: equal to
        # rounding (the angle sums are formed in the same order)
        assert rel(other, out[0]) < 1e-13


def test_ldos_mixed_precision_error_stated(product):
    cfg = CASES["vortex_field"]()
    cfg["numerics"]["norm"] = "l2"
    delta, a = converged_state(product, cfg)
    post = postprocess_config(per_block=-1)
    res = []
    for prec in ("fp64", "mixed"):
        sp = product.Spectroscopy(cfg, post)
        sp.set_order_parameter(0, delta[0])
        sp.set_vector_potential(a[0], a[1])
        for it in range(4):
            sp.compute(prec)
        res.append(sp.ldos())
        sp.close()
    e = rel(res[1], res[0])
    print("mixed-precision LDOS error vs FP64: %.2e" % e)
    assert e < 1e-4


def test_ldos_context_rejects_simulation_calls(product):
    cfg = cases.base(temperature=0.5, angles=8, ppxi=3.0)
    sp = product.Spectroscopy(cfg, postprocess_config(angles=8))
    with pytest.raises(product.SupercongaError):
        product._check(product.lib().sc_compute(sp._h, 0))
    s = product.Simulation(cfg)
    with pytest.raises(product.SupercongaError):
        product._check(product.lib().sc_ldos_compute(s._h, 0, 0))
    sp.close()
    s.close()


def test_postprocess_drop_in_writes_reference_files(product, tmp_path):
    from superconga_amd import io as scio
    cfg = CASES["dwave_hexagon"]()
    cfg["numerics"]["num_iterations_max"] = 3
    run = tmp_path / "run"
    cfg["misc"] = {"save_path": str(run), "save_frequency": -1, "load_path": "",
                   "data_format": "csv", "verbose": False}
    scio.simulate(cfg)
    post = postprocess_config(per_block=3, criterion=1e-6)
    post["misc"] = {"load_path": str(run), "save_path": str(tmp_path / "pp"),
                    "data_format": "csv", "verbose": False}
    converged, ldos = scio.postprocess(post)
    assert converged
    pp = tmp_path / "pp"
    for f in ("ldos.csv", "postprocess_config.json", "postprocess_convergence.csv"):
        assert (pp / f).exists(), f
    rows = (pp / "ldos.csv").read_text().splitlines()
    n0 = ldos.shape[1]
    assert rows[0] == "x_idx,y_idx,e_idx,ldos"
    assert len(rows) == 1 + SPEC["num_energies"] * n0 * n0
    back = np.array([float(r.split(",")[3]) for r in rows[1:]]).reshape(ldos.shape)
    assert np.max(np.abs(back - ldos)) <= 1e-8 * np.max(np.abs(ldos))
    conv = (pp / "postprocess_convergence.csv").read_text().splitlines()
    assert conv[0] == "iteration,boundary_coherence_residual,ldos_residual"
    assert len(conv) > 2


def test_ldos_short_last_block_matches_oracle(product, oracle):
    # 7 energies in blocks of 3: the last block has one energy, so the lane
    # layout and the boundary storage change between blocks.  The block's
    # residual is taken over its own energies in both codes.
    cfg = CASES["dwave_hexagon"]()
    cfg["numerics"]["norm"] = "l2"
    delta, a = converged_state(product, cfg, iterations=3)
    spec = dict(SPEC, num_energies=7)
    post = postprocess_config(per_block=3, spec=spec)
    sp, o = ldos_pair(product, oracle, cfg, post, delta, a)
    conv = sp.run(1, 0, 200, "fp64")
    oconv = o.ldos_run(1, 0, 200)
    st = sp.status()
    assert conv and oconv
    assert st["iterations"] == o.get_int("num_iterations")
    assert st["block_count"] == 1
    e = rel(sp.ldos(), o.ldos())
    print("short last block: %d iterations, rel err %.2e" % (st["iterations"], e))
    assert e < TOL
    sp.close()
