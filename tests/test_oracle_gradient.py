"""The oracle's discrete adjoint of the free energy (no counterpart in the
reference) against what the mathematics fixes: central finite differences of
the oracle's own free energy, step 1e-6, at 64 seeded grid points per case in
FP64 (re and im parts of every component), and the free energy itself equal
to what one iteration of the oracle reports from the same state.

The function differentiated is F(Delta, A) of ComputeFreeEnergy.h:281-414
on the Green functions of one sweep, with the boundary in-values held fixed;
A enters through the Doppler shift of the Riccati steps and the magnetic
term.  Finite differences of a forward sum over ~1e5 sites
resolve a derivative to ~1e-14 / step, so the bound is 2e-6 of the largest
gradient entry at step 1e-6 (measured 1e-7 to 3e-7) and 1e-7 at step 1e-5
(measured 1e-8): a dropped term, a wrong sign or a transposed tap moves
entries by O(1) of the maximum.
"""
import numpy as np
import pytest

import cases
from oracle import pyoracle as po

CASES = {
    # pair-breaking d-wave edges (rotated crystal axes)
    "dwave_hexagon": lambda: cases.base(
        temperature=0.3, angles=16, ppxi=3.0, order_parameter={"dx2-y2": {"critical_temperature": 1.0}},
        rotation=0.125, components=[cases.regular_polygon_unit(6)]),
    # s-wave vortex with external flux and induced vector potential (Doppler shift)
    "vortex_field": lambda: cases.base(
        temperature=0.4, flux=1.5, lam=4.0, angles=16, ppxi=3.0,
        order_parameter={"s": {"critical_temperature": 1.0,
                               "vortices": [{"center_x": 0.0, "center_y": 0.0,
                                             "winding_number": 1}]}}),
    # two components (d + s) on an octagon with a hole
    "d_plus_s_hole": cases.mixed_small,
}


def state(name, iterations=2):
    o = po.Oracle(CASES[name]())
    o.initial_bulk()
    for _ in range(iterations):
        o.compute()
    return o


def sample_points(o, n, seed):
    """n seeded sites: mostly inside, a few outside next to the edge (the
    extension of the order parameter), a few on the edge itself."""
    rng = np.random.default_rng(seed)
    dom = o.domain()
    inside = np.argwhere(dom > 0)
    near = np.zeros_like(dom, dtype=bool)
    for dy in (-2, -1, 0, 1, 2):
        for dx in (-2, -1, 0, 1, 2):
            near |= np.roll(np.roll(dom > 0, dy, 0), dx, 1)
    outside = np.argwhere(near & (dom == 0))
    k_out = n // 8
    pts = [tuple(p) for p in inside[rng.choice(len(inside), n - k_out, replace=False)]]
    pts += [tuple(p) for p in outside[rng.choice(len(outside), k_out, replace=False)]]
    return pts


def fd_errors(o, pts, step):
    f0, grad = o.free_energy_gradient()
    d = o.delta()
    worst = 0.0
    for c in range(o.Nc):
        for (y, x) in pts:
            for unit in (1.0, 1j):
                dd = d[c].copy()
                dd[y, x] += step * unit
                o.set_order_parameter(c, dd)
                fp, _ = o.free_energy_gradient()
                dd[y, x] -= 2 * step * unit
                o.set_order_parameter(c, dd)
                fm, _ = o.free_energy_gradient()
                o.set_order_parameter(c, d[c])
                fd = (fp - fm) / (2 * step)
                an = grad[c][y, x].real if unit == 1.0 else grad[c][y, x].imag
                worst = max(worst, abs(fd - an))
    return worst / np.max(np.abs(grad)), f0, grad


@pytest.mark.parametrize("name", sorted(CASES))
def test_gradient_matches_central_differences(name):
    o = state(name)
    pts = sample_points(o, 64, seed=5)
    err, f0, grad = fd_errors(o, pts, 1e-6)
    print("%s: F %.10f, |grad| max %.3e, FD error %.2e of max" % (
        name, f0, np.max(np.abs(grad)), err))
    assert err < 2e-6
    # a larger step resolves it further (rounding of the differences ~ 1/step)
    err5, _, _ = fd_errors(o, pts[:8], 1e-5)
    assert err5 < 1e-7


def fd_errors_vector_potential(o, lam, pts, step):
    """Central differences of F in A_x and A_y (the units set_vector_potential
    takes; the oracle stores lambda^2 A) against the adjoint's dF/dA."""
    f0, _, ga = o.free_energy_gradient(vector_potential=True)
    a = o.vector_potential() / lam ** 2
    worst = 0.0
    for (y, x) in pts:
        for f in range(2):
            aa = a.copy()
            aa[f, y, x] += step
            o.set_vector_potential(aa[0], aa[1])
            fp, _ = o.free_energy_gradient()
            aa[f, y, x] -= 2 * step
            o.set_vector_potential(aa[0], aa[1])
            fm, _ = o.free_energy_gradient()
            worst = max(worst, abs((fp - fm) / (2 * step) - ga[f, y, x]))
    o.set_vector_potential(a[0], a[1])
    return worst / np.max(np.abs(ga)), ga


@pytest.mark.parametrize("name", ["d_plus_s_hole", "vortex_field"])
def test_vector_potential_gradient_matches_central_differences(name):
    # A enters F through the Doppler shift of every Riccati step (the
    # energy's real part) and through the magnetic term (1 / 2 lambda^2)
    # sum (curl lambda^2 A)^2 h^2, which alone acts far outside the grain.
    # dF/dA is smaller than dF/dDelta, so the rounding of the differences is
    # larger: about 8e-6 of the largest entry at step 1e-6, falling as
    # 1/step to 2e-8 at step 1e-3 (no floor: a dropped term or a wrong
    # factor would leave one of order 1).
    o = state(name)
    lam = CASES[name]()["physics"]["penetration_depth"]
    pts = sample_points(o, 24, seed=7)
    n = o.N0
    far = [(1, 1), (n // 2, 2), (n - 3, n // 3)]
    err6, ga = fd_errors_vector_potential(o, lam, pts + far, 1e-6)
    err3, _ = fd_errors_vector_potential(o, lam, pts[:5] + far, 1e-3)
    print("%s: |dF/dA| max %.3e, FD error %.2e of max at step 1e-6, %.2e at step 1e-3" % (
        name, np.max(np.abs(ga)), err6, err3))
    assert err6 < 3e-5
    assert err3 < 1e-7
    # both routes carry weight: the far sites only through the curl term
    assert min(abs(ga[:, y, x]).max() for (y, x) in far) > 0.0
    # without an induced potential the gradient in A is zero
    o2 = state("dwave_hexagon")
    _, _, ga2 = o2.free_energy_gradient(vector_potential=True)
    assert not np.any(ga2)


def test_gradient_free_energy_is_the_iterations():
    # F from the gradient evaluation equals what compute() reports from the
    # same state, and the evaluation leaves the state untouched
    # (a fresh context, so that compute() takes no accelerator step first)
    src = state("vortex_field")
    cfg = CASES["vortex_field"]()
    o = po.Oracle(cfg)
    o.set_order_parameter(0, src.delta()[0])
    # the oracle stores lambda^2 A and takes A
    a = src.vector_potential() / cfg["physics"]["penetration_depth"] ** 2
    o.set_vector_potential(a[0], a[1])
    f, _ = o.free_energy_gradient()
    f2, _ = o.free_energy_gradient()
    assert f == f2
    o.compute()
    assert abs(f - o.get("free_energy")) <= 1e-12 * abs(f)


def test_gradient_vanishes_where_delta_is_overwritten():
    # sites outside the grain that take the mean of the nearest inside sites
    # (domain_ops.h:11-57) do not enter F
    o = state("dwave_hexagon")
    _, grad = o.free_energy_gradient()
    dom = o.domain()
    near = np.zeros_like(dom, dtype=bool)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            near |= np.roll(np.roll(dom > 0, dy, 0), dx, 1)
    ring = near & (dom == 0)
    assert np.all(grad[0][ring] == 0.0)
    assert np.max(np.abs(grad[0][dom > 0])) > 1e-3


@pytest.mark.parametrize("name", sorted(CASES))
def test_residual_gradient_matches_central_differences(name):
    # R = 1/2 sum_inside |Delta' - Delta|^2 + 1/2 sum |A' - A|^2 of one plain
    # iteration (boundary in-values fixed) and its gradient (J - I)^T (G(u) - u)
    # by the reverse sweep seeded from the gap equation and the current:
    # central differences at 12 seeded sites per component (re and im) and,
    # with a penetration depth, in A_x and A_y there and at two sites far
    # outside.  Measured: 1e-9 to 2e-7 of the largest entry in Delta, 2e-10
    # to 1e-6 in A, at steps 1e-6 and 1e-4.
    o = state(name)
    lam = CASES[name]()["physics"].get("penetration_depth", -1.0)
    r0, g, ga = o.residual_gradient()
    assert r0 > 0.0
    d = o.delta()
    pts = sample_points(o, 12, seed=3)
    gmax = np.max(np.abs(g))
    for step in (1e-6, 1e-4):
        worst = 0.0
        for c in range(o.Nc):
            for (y, x) in pts:
                for unit in (1.0, 1j):
                    dd = d[c].copy()
                    dd[y, x] += step * unit
                    o.set_order_parameter(c, dd)
                    fp, _, _ = o.residual_gradient()
                    dd[y, x] -= 2 * step * unit
                    o.set_order_parameter(c, dd)
                    fm, _, _ = o.residual_gradient()
                    o.set_order_parameter(c, d[c])
                    an = g[c][y, x].real if unit == 1.0 else g[c][y, x].imag
                    worst = max(worst, abs((fp - fm) / (2 * step) - an))
        print("%s: R %.6e, residual gradient in Delta: FD error %.2e of max at step %g" % (
            name, r0, worst / gmax, step))
        assert worst / gmax < 1e-6
    if lam > 0:
        n = o.N0
        pa = pts + [(1, 1), (n // 2, 2)]
        for step in (1e-6, 1e-4):
            err, _ = fd_residual_a(o, lam, pa, step, ga)
            print("%s: residual gradient in A: FD error %.2e of max at step %g" % (name, err, step))
            assert err < 1e-5
    else:
        assert not np.any(ga)


def fd_residual_a(o, lam, pts, step, ga):
    a = o.vector_potential() / lam ** 2
    worst = 0.0
    for (y, x) in pts:
        for f in range(2):
            aa = a.copy()
            aa[f, y, x] += step
            o.set_vector_potential(aa[0], aa[1])
            fp, _, _ = o.residual_gradient()
            aa[f, y, x] -= 2 * step
            o.set_vector_potential(aa[0], aa[1])
            fm, _, _ = o.residual_gradient()
            worst = max(worst, abs((fp - fm) / (2 * step) - ga[f, y, x]))
    o.set_vector_potential(a[0], a[1])
    return worst / np.max(np.abs(ga)), a


def test_residual_is_the_iterations_residual():
    # R from the gradient evaluation is half the squared l2 distance of one
    # plain iteration's answer from the state: the first compute() of a
    # fresh context takes no accelerator step, so its new order parameter
    # and vector potential are G(u)
    src = state("vortex_field")
    cfg = CASES["vortex_field"]()
    o = po.Oracle(cfg)
    o.set_order_parameter(0, src.delta()[0])
    lam = cfg["physics"]["penetration_depth"]
    a = src.vector_potential() / lam ** 2
    o.set_vector_potential(a[0], a[1])
    r, _, _ = o.residual_gradient()
    r2, _, _ = o.residual_gradient()
    assert r == r2
    d0 = o.delta()[0]
    o.compute()
    dom = o.domain() > 0
    dn = o.delta(True)[0]
    an = o.vector_potential(True) / lam ** 2
    expect = 0.5 * np.sum(np.abs(dn - d0)[dom] ** 2) + 0.5 * np.sum((an - a) ** 2)
    assert abs(r - expect) <= 1e-10 * expect
