"""Jacobian-free Newton-Krylov (sc_newton_krylov) against the fixed point
that plain self-consistency iterations converge to.

The fixed point itself is the reference's: the iterations that reach it are
parity-tested against the oracle iteration by iteration, and its free energy
matches the accelerator known answer of tests/accelerators_test.cu:42-150.
Newton's answer must be that same fixed point: the map G it solves u = G(u)
for is one plain FP64 iteration, so its zero is the state every accelerator
converges to.
"""
import json
import os

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def converged_plain(product, cfg, criterion=1e-10):
    c = json.loads(json.dumps(cfg))
    c["numerics"]["convergence_criterion"] = criterion
    c["accelerator"] = {"name": "congacc"}
    s = product.Simulation(c)
    ok = s.run(0, 0, 4000, "fp64")
    r = s.results()
    print("plain fixed point:", r["iterations"], "iterations, residual %.2e" % r["residual"],
          {k: "%.1e" % r[k] for k in r if k.startswith("residual_")})
    assert ok
    return s


def test_newton_krylov_finds_the_fixed_point(product):
    cfg = cases.accelerator_case("picard")
    ref = converged_plain(product, cfg)
    s = product.Simulation(cfg)
    rep = s.newton_krylov(max_newton=20, krylov_dim=20, tol=1e-10)
    print("newton-krylov:", {k: v for k, v in rep.items() if k != "history"},
          "history", ["%.2e" % h for h in rep["history"]])
    assert rep["converged"], rep
    assert rep["history"][-1] < 1e-10
    # residuals decrease monotonically (the line search only accepts descent)
    assert all(b < a for a, b in zip(rep["history"], rep["history"][1:]))
    assert rep["map_evaluations"] >= rep["krylov_iterations"] + rep["newton_iterations"] + 1
    # a plain iteration from Newton's answer leaves it where it is
    s.compute("fp64")
    r = s.results()
    # (the current and free-energy residuals of a first compute are
    # 0.1 x criterion by definition, ComputeCurrent.h:205-212)
    for k in ("residual_order_parameter", "residual_boundary", "residual_vector_potential"):
        assert r[k] < 1e-8, (k, r)
    # both are within their residuals of the fixed point
    assert rel(s.order_parameter(1), ref.order_parameter(1)) < 1e-6
    assert rel(s.current(), ref.current()) < 1e-5
    assert rel(s.vector_potential(1), ref.vector_potential(1)) < 1e-5
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        ka = json.load(f)["accelerator_free_energy_per_area"]
    assert abs(r["free_energy_per_area"] - ka["value"]) < ka["tolerance_rel"] * abs(ka["value"])
    assert abs(r["free_energy_per_area"] - ref.results()["free_energy_per_area"]) < 1e-8 * abs(
        ka["value"])


def test_newton_krylov_without_vector_potential(product):
    # no induced potential: the state is the order parameter and the boundary
    cfg = cases.base(temperature=0.3, angles=16, ppxi=3.0,
                     order_parameter={"dx2-y2": {"critical_temperature": 1.0}},
                     rotation=0.125, components=[cases.regular_polygon_unit(6)])
    ref = converged_plain(product, cfg)
    s = product.Simulation(cfg)
    rep = s.newton_krylov(max_newton=20, krylov_dim=16, tol=1e-10)
    assert rep["converged"], rep
    s.compute("fp64")
    r = s.results()
    for k in ("residual_order_parameter", "residual_boundary"):
        assert r[k] < 1e-8, (k, r)
    assert rel(s.order_parameter(1), ref.order_parameter(1)) < 1e-6


def test_newton_krylov_rejects_bad_arguments(product):
    s = product.Simulation(cases.accelerator_case("picard"))
    for kw in ({"krylov_dim": 0}, {"krylov_dim": 64}, {"tol": 0.0}, {"max_newton": -1},
               {"jvp_precision": "mixed_refine"}):
        args = dict(max_newton=2, krylov_dim=4, tol=1e-8)
        args.update(kw)
        with pytest.raises(product.SupercongaError):
            s.newton_krylov(**args)
    for restarts, eta in ((0, 0.3), (4, 0.0), (4, 1.0)):
        with pytest.raises(product.SupercongaError):
            s.set_newton_options(restarts, eta)


def test_newton_options_keep_the_fixed_point(product):
    # one GMRES cycle per Newton step with a tight forcing cap, and many
    # cycles with a loose one: the same fixed point
    cfg = cases.accelerator_case("picard")
    ref = converged_plain(product, cfg)
    for restarts, eta, fallback, jvp, fix in ((1, 0.01, True, "fp64", False),
                                              (12, 0.5, True, "fp64", False),
                                              (4, 0.3, False, "mixed", False),
                                              (2, 0.3, True, "fp64", True),
                                              (2, 0.3, True, "mixed", True)):
        s = product.Simulation(cfg)
        s.set_newton_options(restarts, eta, fallback, fix)
        rep = s.newton_krylov(max_newton=40, krylov_dim=20, tol=1e-10, jvp_precision=jvp)
        print("restarts %d, eta_max %g:" % (restarts, eta),
              {k: v for k, v in rep.items() if k != "history"})
        assert rep["converged"], rep
        s.compute("fp64")
        assert s.results()["residual_order_parameter"] < 1e-8
        a, b = s.order_parameter(1), ref.order_parameter(1)
        phase = np.angle(np.sum(np.conj(b) * a))
        assert rel(a * np.exp(-1j * phase), b) < 1e-6
        s.close()


def test_newton_krylov_with_fp32_products_finds_the_fp64_fixed_point(product):
    # Jacobian-vector products on the map with FP32 stepping, residuals in
    # FP64: an inexact Newton direction whose error the FP64 residuals
    # correct, so the answer is the FP64 fixed point
    cfg = cases.accelerator_case("picard")
    ref = converged_plain(product, cfg)
    s = product.Simulation(cfg)
    rep = s.newton_krylov(max_newton=30, krylov_dim=20, tol=1e-10, jvp_precision="mixed")
    print("newton-krylov, FP32 products:", {k: v for k, v in rep.items() if k != "history"},
          "history", ["%.2e" % h for h in rep["history"]])
    assert rep["converged"], rep
    s.compute("fp64")
    r = s.results()
    for k in ("residual_order_parameter", "residual_boundary", "residual_vector_potential"):
        assert r[k] < 1e-8, (k, r)
    assert rel(s.order_parameter(1), ref.order_parameter(1)) < 1e-6
    assert rel(s.vector_potential(1), ref.vector_potential(1)) < 1e-5
    assert abs(r["free_energy_per_area"] - ref.results()["free_energy_per_area"]) < 1e-8 * abs(
        r["free_energy_per_area"])


@pytest.mark.parametrize("acc", ["congacc", "anderson", "picard"])
def test_screened_vector_potential_update_keeps_the_fixed_point(product, acc):
    # accelerator.vector_potential_screening hands the accelerator
    # A + P (G_A - A) instead of G_A: the same fixed points, so every
    # accelerator converges to the plain run's state
    def solve(rho):
        c = cases.accelerator_case(acc)
        c["numerics"]["convergence_criterion"] = 1e-10
        c["numerics"]["norm"] = "l2"
        c["accelerator"]["vector_potential_screening"] = rho
        sim = product.Simulation(c)
        ok = sim.run(0, 0, 3000, "fp64")
        return ok, sim
    ok0, plain = solve(0.0)
    ok1, scr = solve(1.0)
    r0, r1 = plain.results(), scr.results()
    print("%s: plain %d iterations, screened %d iterations" % (acc, r0["iterations"],
                                                              r1["iterations"]))
    assert ok1
    if ok0:
        assert rel(scr.order_parameter(0), plain.order_parameter(0)) < 1e-7
        assert rel(scr.vector_potential(0), plain.vector_potential(0)) < 1e-6
        assert abs(r1["free_energy_per_area"] - r0["free_energy_per_area"]) < 1e-9 * abs(
            r0["free_energy_per_area"])
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        ka = json.load(f)["accelerator_free_energy_per_area"]
    assert abs(r1["free_energy_per_area"] - ka["value"]) < ka["tolerance_rel"] * abs(ka["value"])


@pytest.mark.parametrize("jvp", ["fp64", "mixed", "tangent"])
def test_batched_newton_krylov_equals_each_grain_alone(product, jvp):
    # sc_batch_newton_krylov: every grain's own solve, the maps of all of
    # them in shared sweep launches; each grain must end where its solve
    # alone ends (the batch sweep and the single-grain sweep agree to the
    # rounding of their energy sums)
    cfgs = [cases.accelerator_case("picard"),
            cases.base(temperature=0.3, angles=16, ppxi=3.0,
                       order_parameter={"dx2-y2": {"critical_temperature": 1.0}},
                       rotation=0.125, components=[cases.regular_polygon_unit(6)]),
            cases.base(temperature=0.4, flux=0.5, lam=6.0, angles=16, ppxi=3.0, width=8.0)]
    b = product.Batch(cfgs)
    starts = [b.grain(i).order_parameter(0) for i in range(len(cfgs))]
    reps = b.newton_krylov(max_newton=30, krylov_dim=20, tol=1e-10, jvp_precision=jvp)
    for i, cfg in enumerate(cfgs):
        s = product.Simulation(cfg)
        for c, d in enumerate(starts[i]):
            s.set_order_parameter(c, d)
        rep = s.newton_krylov(max_newton=30, krylov_dim=20, tol=1e-10, jvp_precision=jvp)
        print("grain %d: batch %s | alone %s" % (
            i, {k: v for k, v in reps[i].items() if k != "history"},
            {k: v for k, v in rep.items() if k != "history"}))
        assert reps[i]["converged"] and rep["converged"]
        assert rel(b.grain(i).order_parameter(0), s.order_parameter(0)) < 1e-8
        # a plain iteration from the batch's answer leaves it in place
        g = b.grain(i)
        g.compute("fp64")
        r = g.results()
        for k in ("residual_order_parameter", "residual_boundary"):
            assert r[k] < 1e-8, (i, k, r)
        s.close()


def test_batched_newton_krylov_on_selected_grains(product):
    # only the selected grains are solved; the others keep their state
    cfgs = [cases.accelerator_case("picard"), cases.accelerator_case("congacc"),
            cases.base(temperature=0.4, flux=0.5, lam=6.0, angles=16, ppxi=3.0, width=8.0)]
    b = product.Batch(cfgs)
    before = b.grain(1).order_parameter(0)
    reps = b.newton_krylov(max_newton=30, krylov_dim=20, tol=1e-10, grains=[0, 2])
    assert reps[0]["converged"] and reps[2]["converged"]
    assert reps[1]["newton_iterations"] == 0 and reps[1]["map_evaluations"] == 0
    assert np.array_equal(b.grain(1).order_parameter(0), before)
    for i in (0, 2):
        g = b.grain(i)
        g.compute("fp64")
        assert g.results()["residual_order_parameter"] < 1e-8


def test_screening_switched_between_iterations(product):
    # sc_set_screening: on mid-run and off again, the run still ends at the
    # plain run's fixed point
    cfg = cases.accelerator_case("picard")
    cfg["numerics"]["convergence_criterion"] = 1e-10
    cfg["numerics"]["norm"] = "l2"
    plain = product.Simulation(cfg)
    assert plain.run(0, 0, 3000, "fp64")
    s = product.Simulation(cfg)
    for _ in range(5):
        s.compute("fp64")
    s.set_screening(0.9)
    for _ in range(20):
        s.compute("fp64")
    s.set_screening(0.0)
    assert s.run(0, 0, 3000, "fp64")
    assert rel(s.order_parameter(0), plain.order_parameter(0)) < 1e-7
    assert rel(s.vector_potential(0), plain.vector_potential(0)) < 1e-6
    with pytest.raises(product.SupercongaError):
        s.set_screening(-1.0)


def test_newton_krylov_with_tangent_products(product):
    # exact products: the same fixed point as the finite differences, with
    # map evaluations only for the Newton points and the line search (each
    # product is a tangent-linear sweep, not a map evaluation)
    for cfg in (cases.accelerator_case("picard"),
                cases.base(temperature=0.4, flux=0.5, lam=6.0, angles=16, ppxi=3.0, width=8.0)):
        ref = converged_plain(product, cfg)
        fd = product.Simulation(cfg)
        rep_fd = fd.newton_krylov(max_newton=20, krylov_dim=20, tol=1e-10)
        s = product.Simulation(cfg)
        rep = s.newton_krylov(max_newton=20, krylov_dim=20, tol=1e-10, jvp_precision="tangent")
        print("tangent products:", {k: v for k, v in rep.items() if k != "history"},
              "| differences:", {k: v for k, v in rep_fd.items() if k != "history"})
        assert rep["converged"], rep
        assert rep["map_evaluations"] < rep["newton_iterations"] + 8
        assert rep_fd["map_evaluations"] > rep_fd["krylov_iterations"]
        s.compute("fp64")
        r = s.results()
        for k in ("residual_order_parameter", "residual_boundary"):
            assert r[k] < 1e-8, (k, r)
        # the global phase of the order parameter is a symmetry of the map
        # (a circle of fixed points): compare after removing it
        a, b = s.order_parameter(1), ref.order_parameter(1)
        phase = np.angle(np.sum(np.conj(b) * a))
        print("global phase offset from the plain run: %.2e rad" % phase)
        assert rel(a * np.exp(-1j * phase), b) < 1e-6
        assert abs(r["free_energy"] - ref.results()["free_energy"]) < 1e-8 * abs(
            r["free_energy"])
        s.close()
        fd.close()
