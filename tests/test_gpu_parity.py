"""Parity of the HIP self-consistency path with the CPU oracle, through the
C ABI (superconga_amd wraps include/superconga_amd.h with ctypes).

Bar (FP64): every grid point of order parameter, current, vector potential,
free-energy density and boundary storage within 1e-10 of the oracle,
relative to the field's largest magnitude:
    max |gpu - oracle| <= 1e-10 * max |oracle|
Scalars (free energy) within 1e-10 relative.  Residuals are normalised
differences of consecutive iterates, |x_new - x_old| / |x_old|, so their
rounding error is that of x (1e-16 relative) divided by the residual itself:
they are compared as |r_gpu - r_oracle| <= 1e-10 |r_oracle| + 1e-13.  Mixed precision
(FP32 stepping, FP64 sums) states its error against the FP64 oracle and is
bounded by MIXED_TOL.
"""
import json
import math
import os

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

TOL = 1e-10
MIXED_TOL = 1e-4
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    scale = np.max(np.abs(b))
    if scale == 0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / scale)


def srel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def pair(product, oracle, cfg):
    s = product.Simulation(cfg)
    o = oracle.Oracle(cfg)
    o.initial_bulk()
    for c, d in enumerate(o.delta()):
        s.set_order_parameter(c, d)
    return s, o


def lam2(cfg):
    lam = cfg["physics"].get("penetration_depth", -1.0)
    return lam * lam if lam > 0 else 1.0


def compare_state(s, o, cfg, tol=TOL, what=""):
    errs = {
        "delta_result": rel(s.order_parameter(1), o.delta(True)),
        "delta_iterate": rel(s.order_parameter(0), o.delta(False)),
        "current": rel(s.current(), o.current()),
        "free_energy_density": rel(
            s.free_energy_density(),
            o.array("free_energy_grid", 2 * o.N0 ** 2)[:o.N0 ** 2].reshape(o.N0, o.N0)),
    }
    if cfg["physics"].get("penetration_depth", -1.0) > 0:
        errs["vector_potential"] = rel(s.vector_potential(1),
                                       o.vector_potential(True) / lam2(cfg))
        errs["vector_potential_iterate"] = rel(s.vector_potential(0),
                                               o.vector_potential(False) / lam2(cfg))
    r = s.results()
    for k in ("free_energy", "free_energy_per_area", "magnetic_free_energy"):
        errs[k] = srel(r[k], o.get(k))
    for k in ("residual_order_parameter", "residual_current", "residual_boundary"):
        # relative error with the absolute floor 1e-13 folded in
        errs[k] = abs(r[k] - o.get(k)) / (abs(o.get(k)) + 1e-13 / tol)
    for k in ("residual_free_energy", "residual"):
        # The free-energy residual |F_n - F_n-1| / |F_n-1| inherits the
        # absolute error of F / |F| (F is a difference of large sums that the
        # two codes add in different orders, equal to ~1e-11 relative on the
        # first iterations), so it is compared in absolute terms; so is the
        # overall residual, of which it can be the maximum.
        errs[k] = abs(r[k] - o.get(k)) / (1.0 + abs(o.get(k)))
    bad = {k: v for k, v in errs.items() if not v <= tol}
    print(what, {k: "%.2e" % v for k, v in errs.items()}, flush=True)
    assert not bad, (what, bad)


def compare_boundary(s, o, tol=TOL):
    for io in (0, 1):
        for d in (0, 1):
            assert rel(s.boundary(io, d), o.boundary(io, d)) <= tol, (io, d)


# ---------------- unit-level device arithmetic ----------------

def test_riccati_steps_match_oracle(product, oracle):
    rng = np.random.default_rng(11)
    n = 257
    d = rng.normal(size=n) + 1j * rng.normal(size=n)
    e = 0.3 * rng.normal(size=n) + 1j * (np.abs(rng.normal(size=n)) + 0.01)
    y0 = 0.4 * (rng.normal(size=n) + 1j * rng.normal(size=n))
    h = rng.uniform(0.0, 0.6, size=n)
    ref = np.array([oracle.riccati_step(d[i], e[i], y0[i], h[i]) for i in range(n)])
    gpu = product.riccati_steps(d, e, y0, h)
    assert np.max(np.abs(gpu - ref) / np.maximum(np.abs(ref), 1e-3)) < 1e-13
    mixed = product.riccati_steps(d, e, y0, h, precision="mixed")
    err = np.max(np.abs(mixed - ref) / np.maximum(np.abs(ref), 1e-3))
    print("mixed riccati step error vs FP64 oracle: %.2e" % err)
    assert err < 1e-5


def test_riccati_steps_closed_forms(product, oracle):
    # tests/riccati_solutions_test.cu:21-177 on the device.
    d, e = complex(1, 0.5), complex(0.5, 1)
    g, gt = oracle.homogeneous(d, e)
    out = product.riccati_steps([d, -d.conjugate()], e, [g, gt], 1.0)
    assert abs(out[0] - g) < 1e-14 and abs(out[1] - gt) < 1e-14
    y0 = complex(0, 0.5)
    out = product.riccati_steps([1.0, -1.0], 1j, y0, 0.0)
    assert np.all(out == y0)


def test_poisson_matches_oracle(product, oracle):
    rng = np.random.default_rng(2)
    for n, h, ng in ((17, 0.25, 150), (64, 0.1, 2000)):
        j = rng.normal(size=(2, n, n))
        a = product.poisson_solve(j, h, ng)
        b = oracle.vector_potential(n, h, ng, j)
        assert rel(a, b) < 1e-12, n


@pytest.mark.parametrize("max_error", [1e-3, 1e-6])
def test_poisson_filters_match_oracle(product, oracle, max_error):
    # the eigen-truncated kernel of vector_potential_error > 0
    rng = np.random.default_rng(5)
    for n, h, ng in ((33, 0.2, 700), (120, 0.05, 9000)):
        j = rng.normal(size=(2, n, n))
        a = product.poisson_solve(j, h, ng, max_error=max_error)
        b = oracle.vector_potential(n, h, ng, j, max_error=max_error)
        exact = product.poisson_solve(j, h, ng)
        assert rel(a, b) < 1e-12, n
        # the truncation is visible (not the exact kernel) and bounded
        assert 0 < rel(a, exact) < 100 * max_error


def test_poisson_analytic_disc(product):
    # tests/physics/vector_potential_test.cu:126-226 (rel Frobenius 2e-4).
    n, h = 256, 0.1
    c = 0.5 * (n - 1)
    R = h * (n - 1) * 0.4
    idx = np.arange(n)
    y = h * (idx[:, None] - c)
    x = h * (idx[None, :] - c)
    r = np.hypot(x, y)
    inside = r <= R
    j = np.stack([np.where(inside, -y, 0.0), np.where(inside, x, 0.0)])
    with np.errstate(divide="ignore"):
        ang = np.where(inside, 0.125 * (2 * R * R - r * r), 0.125 * R ** 4 / (r * r))
    expected = np.stack([-ang * y, ang * x])
    a = product.poisson_solve(j, h, int(inside.sum()))
    assert math.sqrt(np.sum((a - expected) ** 2) / np.sum(expected ** 2)) < 2e-4


# ---------------- whole iterations ----------------

ITERATION_CASES = {
    "disc_field_picard": cases.accelerator_case("picard"),
    "disc_field_congacc": cases.accelerator_case("congacc"),
    "disc_field_anderson": cases.accelerator_case("anderson"),
    "disc_field_bb": cases.accelerator_case("bb"),
    "disc_field_polyak": cases.accelerator_case("polyak"),
    "disc_field_broyden": cases.accelerator_case("broyden"),
    "disc_field_broyden_rank3": cases.accelerator_case("broyden", rank=3, step_size=0.8),
    "octagon_hole_d_plus_s": cases.mixed_small(),
    "square_dxy_landau": cases.base(
        order_parameter={"dxy": {"critical_temperature": 1.0}}, temperature=0.4,
        flux=0.5, lam=-1.0, angles=20, ppxi=4.0, gauge="landau",
        components=[cases.regular_polygon_unit(4)], accelerator="anderson", rank=3),
    # angle counts off the Boole rule (N % 3, N % 2 and plain weights,
    # FermiSurface.h:119-138) and a high temperature with a low cutoff (two
    # Ozaki poles)
    "disc_7_angles_hot": cases.base(temperature=0.95, angles=7, ppxi=3.0, flux=0.5,
                                    cutoff=3.0),
    "square_10_angles": cases.base(
        order_parameter={"dx2-y2": {"critical_temperature": 1.0}}, temperature=0.4,
        angles=10, ppxi=3.0, rotation=0.1, components=[cases.regular_polygon_unit(4)]),
    "hexagon_18_angles_field": cases.base(
        temperature=0.5, angles=18, ppxi=3.0, flux=1.0, lam=6.0,
        components=[cases.regular_polygon_unit(6)]),
    "polygon_g_solenoid": cases.base(
        order_parameter={"g": {"critical_temperature": 1.0}}, temperature=0.3,
        flux=1.5, lam=3.0, angles=16, ppxi=3.0, gauge="solenoid", width=12.0,
        components=[{"polygon": {"vertices_x": [-0.5, 0.5, 0.3, -0.4],
                                 "vertices_y": [-0.45, -0.5, 0.5, 0.2]}}]),
}


@pytest.mark.parametrize("name", sorted(ITERATION_CASES))
def test_iterations_match_oracle(product, oracle, name):
    cfg = ITERATION_CASES[name]
    s, o = pair(product, oracle, cfg)
    for it in range(4):
        s.compute("fp64")
        o.compute()
        compare_state(s, o, cfg, what="%s it %d" % (name, it))
    compare_boundary(s, o)


def test_burnin_then_iterations(product, oracle):
    cfg = cases.mixed_small()
    s, o = pair(product, oracle, cfg)
    for _ in range(2):
        s.burnin("fp64")
        o.burnin()
    compare_boundary(s, o)
    for it in range(2):
        s.compute("fp64")
        o.compute()
        compare_state(s, o, cfg, what="burnin+compute %d" % it)


def test_sweep_shapes_agree(product, oracle):
    # 10 Ozaki energies: 1, 2 and 4 energies per lane give 10, 5 and 3
    # groups per column (lanes past the last whole column idle, the groups
    # summed lane by lane), 7 and 8 give two groups, 14 and 16 one
    cfg = cases.mixed_small()
    out = []
    for el, waves, slots in ((1, 4, 16), (2, 4, 16), (4, 4, 16), (7, 2, 8), (8, 1, 4),
                             (14, 2, 8), (16, 1, 2)):
        s, o = pair(product, oracle, cfg)
        s.set_sweep_shape(el, waves, slots)
        s.compute("fp64")
        out.append((s.order_parameter(1), s.current()))
    o.compute()
    for d, j in out:
        assert rel(d, o.delta(True)) <= TOL
        assert rel(j, o.current()) <= TOL


def test_autotune_leaves_state_and_parity(product, oracle):
    cfg = cases.mixed_small()
    s, o = pair(product, oracle, cfg)
    s.compute("fp64")
    o.compute()
    for prec in ("fp64", "mixed"):
        shape = s.autotune(prec)
        print("autotuned", prec, shape)
        assert len(shape) == 3
    for it in range(2):
        s.compute("fp64")
        o.compute()
        compare_state(s, o, cfg, what="after autotune %d" % it)


def test_mixed_shapes_agree(product):
    cfg = cases.mixed_small()
    out = []
    for el, waves, slots in ((4, 2, 8), (7, 1, 4), (8, 2, 8), (16, 2, 8)):
        s = product.Simulation(cfg)
        s.set_sweep_shape(el, waves, slots, precision="mixed")
        s.compute("mixed")
        out.append(s.order_parameter(1))
    for d in out[1:]:
        # FP32 stepping: the compiled variants round differently (their
        # multiply-adds are contracted differently), so they agree to FP32
        # rounding, not bit for bit.
        assert rel(d, out[0]) < 1e-6


@pytest.mark.parametrize("accelerator", ["congacc", "broyden"])
def test_run_to_convergence_known_answer(product, oracle, accelerator):
    # tests/accelerators_test.cu:142: every accelerator converges to the same
    # free energy (the reference checks its five; Broyden is held to it too)
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        ka = json.load(f)["accelerator_free_energy_per_area"]
    cfg = cases.accelerator_case(accelerator)
    s, o = pair(product, oracle, cfg)
    conv = s.run(0, 0, 100, "fp64")
    oconv = o.run(0, 0, 100)
    r = s.results()
    assert conv and oconv
    assert r["iterations"] == o.get_int("num_iterations")
    assert srel(r["free_energy_per_area"], o.get("free_energy_per_area")) < 1e-9
    assert abs(r["free_energy_per_area"] - ka["value"]) < ka["tolerance_rel"] * abs(ka["value"])


def test_mixed_precision_error_stated(product, oracle):
    cfg = cases.mixed_small()
    s, o = pair(product, oracle, cfg)
    for it in range(3):
        s.compute("mixed")
        o.compute()
    err = {
        "delta": rel(s.order_parameter(1), o.delta(True)),
        "current": rel(s.current(), o.current()),
        "vector_potential": rel(s.vector_potential(1), o.vector_potential(True) / lam2(cfg)),
        "free_energy": srel(s.results()["free_energy"], o.get("free_energy")),
    }
    print("mixed precision (FP32 stepping, FP64 sums) error vs FP64 oracle:",
          {k: "%.2e" % v for k, v in err.items()})
    assert all(v < MIXED_TOL for v in err.values()), err


def test_boundary_storage_roundtrip(product, oracle):
    cfg = cases.accelerator_case("picard")
    s, o = pair(product, oracle, cfg)
    s.compute("fp64")
    for d in (0, 1):
        v = s.boundary(0, d)
        w = v * (1 + 0.5j)
        s.set_boundary(d, w)
        assert np.array_equal(s.boundary(0, d), w)


def test_order_parameter_outside_grain_is_masked(product, oracle):
    cfg = cases.accelerator_case("picard")
    s, o = pair(product, oracle, cfg)
    dom = o.domain().astype(bool)
    s.set_order_parameter(0, np.ones((s.N0, s.N0), dtype=complex))
    got = s.order_parameter(0)[0]
    assert np.all(got[~dom] == 0) and np.all(got[dom] == 1)


# ---------------- full size: the d-wave octagon ----------------

def test_octagon_full_size_invariants(product):
    cfg = cases.octagon_dwave()
    s = product.Simulation(cfg)
    assert (s.N0, s.N1, s.Ne, s.Na) == (930, 1316, 28, 128)
    s.compute("fp64")
    d = s.order_parameter(1)[0]
    j = s.current()
    r = s.results()
    # Current conservation over the grain.
    ng = s.dims.num_grain_sites
    jabs = np.abs(j).sum() / ng
    mean = (j[0].sum() / ng, j[1].sum() / ng)
    # The octagon and the 45-degree d-wave basis are symmetric under a
    # 90-degree rotation: |Delta| maps onto itself (up to the single-precision
    # rotation angles of the interpolation, as in the reference).
    a = np.abs(d)
    sym = rel(np.rot90(a), a)
    print("octagon: mean j %s, mean |j| %.3e, rot90 asymmetry %.2e, sweep %.2f ms"
          % (mean, jabs, sym, r["last_sweep_ms"]))
    assert abs(mean[0]) < 1e-6 * jabs + 1e-14
    assert abs(mean[1]) < 1e-6 * jabs + 1e-14
    assert sym < 1e-4
    assert np.isfinite(r["free_energy_per_area"]) and r["free_energy_per_area"] < 0
    assert r["last_sweep_ms"] > 0


def test_mixed_refine_reaches_the_fp64_fixed_point(product):
    # SC_PRECISION_MIXED_REFINE: FP32 stepping until the residual reaches the
    # FP32 floor, then FP64 until the criterion; the converged state must be
    # the FP64 run's fixed point to the criterion's accuracy
    from superconga_amd import workloads
    cfg = workloads.swave_disc()
    cfg["numerics"]["convergence_criterion"] = 1e-8
    ref = product.Simulation(cfg)
    d0 = ref.order_parameter(0)[0]
    assert ref.run(0, 0, 3000, "fp64")
    mix = product.Simulation(cfg)
    mix.set_order_parameter(0, d0)
    assert mix.run(0, 0, 3000, "mixed_refine")
    r64, rm = ref.results(), mix.results()
    print("fp64: %d iterations; mixed refine: %d iterations, %d of them FP32"
          % (r64["iterations"], rm["iterations"], rm["mixed_iterations"]), flush=True)
    assert rm["mixed_iterations"] > 0
    assert rm["converged"] == 1 and rm["residual"] < 1e-8
    # The zero-field disc carries only the weak currents that the seeded
    # noise of its initial state leaves (far below the natural unit j0), so
    # the current's difference is measured against the order parameter's
    # scale (both fields are in the reference's natural units, O(1) in a
    # superconductor) rather than against that noise.
    jref = ref.current()
    dscale = float(np.max(np.abs(ref.order_parameter(0))))
    print("max |j| %.3e, max |Delta| %.3e" % (np.max(np.abs(jref)), dscale))
    err = {
        "delta": rel(mix.order_parameter(0), ref.order_parameter(0)),
        "current": float(np.max(np.abs(mix.current() - jref))
                         / max(float(np.max(np.abs(jref))), dscale)),
        "free_energy": srel(rm["free_energy"], r64["free_energy"]),
    }
    print("mixed refine vs FP64 fixed point:", {k: "%.2e" % v for k, v in err.items()})
    assert all(v < 1e-8 for v in err.values()), err
