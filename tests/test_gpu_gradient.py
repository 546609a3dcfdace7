"""The free-energy gradient on the GPU (sc_free_energy_gradient_ex, the
discrete adjoint of one iteration) against the oracle's adjoint, which
tests/test_oracle_gradient.py pins by central differences: every grid point
of dF/dRe Delta, dF/dIm Delta and dF/dA within 1e-10 of the largest entry, F
within 1e-12 relative, in FP64.  The same gradients through torch.autograd,
and central-difference checks of the product itself on a two-component
d + is square, in Delta and in A.
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

TOL = 1e-10

CASES = {
    "dwave_hexagon": lambda: cases.base(
        temperature=0.3, angles=16, ppxi=3.0,
        order_parameter={"dx2-y2": {"critical_temperature": 1.0}}, rotation=0.125,
        components=[cases.regular_polygon_unit(6)]),
    "vortex_field": lambda: cases.base(
        temperature=0.4, flux=1.5, lam=4.0, angles=16, ppxi=3.0,
        order_parameter={"s": {"critical_temperature": 1.0,
                               "vortices": [{"center_x": 0.0, "center_y": 0.0,
                                             "winding_number": 1}]}}),
    "d_plus_s_hole": cases.mixed_small,
    # odd angle count (plain weights) and a polygon with g-wave pairing
    "polygon_g_7_angles": lambda: cases.base(
        order_parameter={"g": {"critical_temperature": 1.0}}, temperature=0.3,
        flux=1.0, lam=3.0, angles=7, ppxi=3.0, gauge="solenoid", width=12.0,
        components=[{"polygon": {"vertices_x": [-0.5, 0.5, 0.3, -0.4],
                                 "vertices_y": [-0.45, -0.5, 0.5, 0.2]}}]),
}


def lam(cfg):
    return cfg["physics"].get("penetration_depth", -1.0)


def state(product, cfg, iterations=2):
    """An order parameter and vector potential a few iterations in."""
    s = product.Simulation(cfg)
    for _ in range(iterations):
        s.compute("fp64")
    d = s.order_parameter(0)
    a = s.vector_potential(0)
    s.close()
    return d, a


def fresh_pair(product, oracle, cfg, d, a):
    """A new product grain and a new oracle context on the same state (both
    start their boundary in-values from it)."""
    s = product.Simulation(cfg)
    o = oracle.Oracle(cfg)
    for c in range(d.shape[0]):
        s.set_order_parameter(c, d[c])
        o.set_order_parameter(c, d[c])
    if lam(cfg) > 0:
        s.set_vector_potential(a[0], a[1])
        o.set_vector_potential(a[0], a[1])
    return s, o


@pytest.mark.parametrize("name", sorted(CASES))
def test_gradient_matches_oracle(product, oracle, name):
    cfg = CASES[name]()
    d, a = state(product, cfg)
    s, o = fresh_pair(product, oracle, cfg, d, a)
    f, g, ga = s.free_energy_gradient(vector_potential=True)
    fo, go, gao = o.free_energy_gradient(vector_potential=True)
    err = float(np.max(np.abs(g - go)) / np.max(np.abs(go)))
    if lam(cfg) > 0:
        err_a = float(np.max(np.abs(ga - gao)) / np.max(np.abs(gao)))
    else:
        assert not np.any(ga) and not np.any(gao)
        err_a = 0.0
    print("%s: F %.12f (oracle %.12f), gradient error %.2e of max, in A %.2e of max" % (
        name, f, fo, err, err_a))
    assert abs(f - fo) <= 1e-12 * abs(fo)
    assert err < TOL
    assert err_a < TOL
    # the evaluation leaves the state as it was: the next iteration matches
    # the oracle's
    s.compute("fp64")
    o.compute()
    assert abs(s.results()["free_energy"] - o.get("free_energy")) <= 1e-10 * abs(fo)
    dd = s.order_parameter(1)
    assert float(np.max(np.abs(dd - o.delta(True))) / np.max(np.abs(dd))) < TOL
    s.close()


def test_gradient_is_the_iterations_free_energy(product):
    cfg = CASES["vortex_field"]()
    d, a = state(product, cfg)
    s = product.Simulation(cfg)
    s.set_order_parameter(0, d[0])
    s.set_vector_potential(a[0], a[1])
    f, _ = s.free_energy_gradient()
    f2, _ = s.free_energy_gradient()
    assert abs(f - f2) <= 1e-14 * abs(f)
    s.compute("fp64")
    assert abs(f - s.results()["free_energy"]) <= 1e-12 * abs(f)
    s.close()


def test_autograd_backward_is_the_adjoint(product):
    torch = pytest.importorskip("torch")
    from superconga_amd.autograd import free_energy
    cfg = CASES["d_plus_s_hole"]()
    d, a = state(product, cfg)
    s = product.Simulation(cfg)
    s.set_vector_potential(a[0], a[1])
    planes = np.stack([d.real, d.imag], axis=1)
    x = torch.tensor(planes, dtype=torch.float64, requires_grad=True)
    f = free_energy(s, x)
    loss = 2.0 * f + (x * x).sum()
    loss.backward()
    f0, g = s.free_energy_gradient()
    gp = np.stack([g.real, g.imag], axis=1)
    assert abs(float(f.detach()) - f0) <= 1e-14 * abs(f0)
    expect = 2.0 * gp + 2.0 * planes
    assert float(np.max(np.abs(x.grad.numpy() - expect))) <= 1e-14 * float(np.max(np.abs(expect)))
    # a complex tensor through view_as_real.  Each evaluation is a new run
    # of the adjoint, whose FP64 atomic scatters add in a different order
    # every time: evaluations agree to rounding, not bit for bit.
    z = torch.tensor(d, dtype=torch.complex128, requires_grad=True)
    fz = free_energy(s, torch.view_as_real(z).movedim(-1, 1))
    fz.backward()
    assert np.allclose(z.grad.numpy().real, g.real, rtol=0, atol=1e-14 * np.max(np.abs(g)))
    assert np.allclose(z.grad.numpy().imag, g.imag, rtol=0, atol=1e-14 * np.max(np.abs(g)))
    s.close()


def test_gradient_central_differences_d_plus_is(product):
    # a two-component d + is square in a field: central differences of the
    # product's own F at 64 seeded sites, step 1e-5, against its adjoint
    cfg = cases.base(temperature=0.3, angles=32, ppxi=4.0, flux=1.0, lam=-1.0,
                     width=12.0, components=[cases.regular_polygon_unit(4)],
                     order_parameter={"dx2-y2": {"critical_temperature": 1.0},
                                      "s": {"critical_temperature": 0.5}})
    s = product.Simulation(cfg)
    s.compute("fp64")
    d = s.order_parameter(0)
    d[1] = 0.2j * d[0]
    s.close()
    s = product.Simulation(cfg)
    for c in range(2):
        s.set_order_parameter(c, d[c])
    f0, g = s.free_energy_gradient()
    dom = s.domain() > 0
    rng = np.random.default_rng(11)
    pts = np.argwhere(dom)[rng.choice(int(dom.sum()), 64, replace=False)]
    step = 1e-5
    worst = 0.0
    for (y, x) in pts:
        for c in range(2):
            for unit in (1.0, 1j):
                dd = d[c].copy()
                dd[y, x] += step * unit
                s.set_order_parameter(c, dd)
                fp, _ = s.free_energy_gradient()
                dd[y, x] -= 2 * step * unit
                s.set_order_parameter(c, dd)
                fm, _ = s.free_energy_gradient()
                s.set_order_parameter(c, d[c])
                fd = (fp - fm) / (2 * step)
                an = g[c][y, x].real if unit == 1.0 else g[c][y, x].imag
                worst = max(worst, abs(fd - an))
    err = worst / float(np.max(np.abs(g)))
    print("d+is square: F %.10f, |grad| max %.3e, FD error %.2e of max" % (
        f0, np.max(np.abs(g)), err))
    assert err < 1e-6
    s.close()


def test_gradient_central_differences_full_size_d_plus_is(product):
    # the two-component d + is square at its full size (side 30 xi0 at 20
    # points per xi0, 128 angles, one flux quantum, penetration depth 3 xi0)
    # two iterations in: central differences of F at 64 seeded inside sites,
    # both components, real and imaginary parts, at steps 1e-6 and 1e-5
    import time
    from superconga_amd import workloads
    cfg = workloads.dplus_is_square()
    s = product.Simulation(cfg)
    d0 = s.order_parameter(0)
    for c, x in enumerate(workloads.d_plus_is_initial(d0[0])):
        s.set_order_parameter(c, x)
    for _ in range(2):
        s.compute("fp64")
    d = s.order_parameter(0)
    t = time.perf_counter()
    f0, g = s.free_energy_gradient()
    t = time.perf_counter() - t
    f1, _ = s.free_energy_gradient()
    dom = s.domain() > 0
    rng = np.random.default_rng(5)
    pts = np.argwhere(dom)[rng.choice(int(dom.sum()), 64, replace=False)]
    gmax = float(np.max(np.abs(g)))
    err = {}
    for step in (1e-6, 1e-5):
        worst = 0.0
        for (y, x) in pts:
            for c in range(2):
                for unit in (1.0, 1j):
                    dd = d[c].copy()
                    dd[y, x] += step * unit
                    s.set_order_parameter(c, dd)
                    fp, _ = s.free_energy_gradient()
                    dd[y, x] -= 2 * step * unit
                    s.set_order_parameter(c, dd)
                    fm, _ = s.free_energy_gradient()
                    s.set_order_parameter(c, d[c])
                    fd = (fp - fm) / (2 * step)
                    an = g[c][y, x].real if unit == 1.0 else g[c][y, x].imag
                    worst = max(worst, abs(fd - an))
        err[step] = worst / gmax
    print("full-size d+is square: N0 %d, F %.10f (again: %+.1e), |grad| max %.3e, one gradient "
          "%.1f ms, FD error of max: %.2e at step 1e-6, %.2e at step 1e-5" % (
              s.N0, f0, f1 - f0, gmax, 1e3 * t, err[1e-6], err[1e-5]))
    # a step changes F (about -17, a sum over 3.6e5 sites of terms far
    # larger than the change) by about 1e-10 of itself: the rounding of F
    # divided by the step is about 1e-5 of the largest entry (of order
    # h^2 |Delta|) at step 1e-6 and shrinks as 1/step, which a wrong
    # gradient would not
    assert err[1e-6] < 2e-4
    assert err[1e-5] < 2e-5
    s.close()


def test_autograd_backward_in_the_vector_potential(product):
    torch = pytest.importorskip("torch")
    from superconga_amd.autograd import free_energy
    cfg = CASES["vortex_field"]()
    d, a = state(product, cfg)
    s = product.Simulation(cfg)
    planes = np.stack([d.real, d.imag], axis=1)
    x = torch.tensor(planes, dtype=torch.float64, requires_grad=True)
    at = torch.tensor(a, dtype=torch.float64, requires_grad=True)
    f = free_energy(s, x, at)
    (3.0 * f).backward()
    f0, g, ga = s.free_energy_gradient(vector_potential=True)
    assert abs(float(f.detach()) - f0) <= 1e-14 * abs(f0)
    assert np.allclose(at.grad.numpy(), 3.0 * ga, rtol=0, atol=1e-14 * np.max(np.abs(3.0 * ga)))
    assert np.allclose(x.grad.numpy(), 3.0 * np.stack([g.real, g.imag], axis=1), rtol=0,
                       atol=1e-14 * 3.0 * np.max(np.abs(g)))
    s.close()


def test_vector_potential_gradient_central_differences_full_size_d_plus_is(product):
    # the full-size d + is square two iterations in: central differences of
    # F in A_x and A_y at 64 seeded sites (inside and near the edge) and
    # three far outside, where only the magnetic term acts, at steps 1e-6
    # and 1e-3
    from superconga_amd import workloads
    cfg = workloads.dplus_is_square()
    s = product.Simulation(cfg)
    d0 = s.order_parameter(0)
    for c, x in enumerate(workloads.d_plus_is_initial(d0[0])):
        s.set_order_parameter(c, x)
    for _ in range(2):
        s.compute("fp64")
    a = s.vector_potential(0)
    s.set_vector_potential(a[0], a[1])
    f0, _, ga = s.free_energy_gradient(vector_potential=True)
    dom = s.domain() > 0
    rng = np.random.default_rng(9)
    pts = [tuple(p) for p in np.argwhere(dom)[rng.choice(int(dom.sum()), 60, replace=False)]]
    n = s.N0
    pts += [(1, 1), (n // 2, 2), (n - 3, n // 3), (n // 2, n - 2)]
    gmax = float(np.max(np.abs(ga)))
    err = {}
    for step, sub in ((1e-6, pts), (1e-3, pts[:12] + pts[-4:])):
        worst = 0.0
        for (y, x) in sub:
            for f in range(2):
                aa = a.copy()
                aa[f, y, x] += step
                s.set_vector_potential(aa[0], aa[1])
                fp, _ = s.free_energy_gradient()
                aa[f, y, x] -= 2 * step
                s.set_vector_potential(aa[0], aa[1])
                fm, _ = s.free_energy_gradient()
                worst = max(worst, abs((fp - fm) / (2 * step) - ga[f, y, x]))
        s.set_vector_potential(a[0], a[1])
        err[step] = worst / gmax
    print("full-size d+is square: |dF/dA| max %.3e, FD error of max: %.2e at step 1e-6, "
          "%.2e at step 1e-3" % (gmax, err[1e-6], err[1e-3]))
    # the rounding of F over the step (5e-5 of the largest entry at 1e-6)
    # falls as 1/step; no floor above 4e-8
    assert err[1e-6] < 3e-4
    assert err[1e-3] < 5e-7
    s.close()


@pytest.mark.parametrize("name", sorted(CASES))
def test_residual_gradient_matches_oracle(product, oracle, name):
    # sc_residual_gradient against the oracle's restatement (pinned by
    # central differences in tests/test_oracle_gradient.py): R and every grid
    # point of the gradient in Delta and A; the evaluation leaves the state
    # as it was (the next iteration still matches the oracle's)
    cfg = CASES[name]()
    d, a = state(product, cfg)
    s, o = fresh_pair(product, oracle, cfg, d, a)
    r, g, ga = s.residual_gradient()
    ro, go, gao = o.residual_gradient()
    err = float(np.max(np.abs(g - go)) / np.max(np.abs(go)))
    err_a = float(np.max(np.abs(ga - gao)) / np.max(np.abs(gao))) if lam(cfg) > 0 else 0.0
    print("%s: R %.12e (oracle %.12e), gradient error %.2e of max, in A %.2e of max" % (
        name, r, ro, err, err_a))
    assert abs(r - ro) <= 1e-10 * ro
    assert err < TOL
    assert err_a < TOL
    if lam(cfg) <= 0:
        assert not np.any(ga)
    s.compute("fp64")
    o.compute()
    dd = s.order_parameter(1)
    assert float(np.max(np.abs(dd - o.delta(True))) / np.max(np.abs(dd))) < TOL
    assert abs(s.results()["free_energy"] - o.get("free_energy")) <= 1e-10 * abs(
        o.get("free_energy"))
    s.close()


def test_residual_gradient_central_differences_full_size_d_plus_is(product):
    # the full-size d + is square two iterations in: central differences of
    # R in Delta (24 seeded inside sites, both components, re and im) and in
    # A (16 sites, two of them far outside) at steps 1e-6 and 1e-4
    import time
    from superconga_amd import workloads
    cfg = workloads.dplus_is_square()
    s = product.Simulation(cfg)
    d0 = s.order_parameter(0)
    for c, x in enumerate(workloads.d_plus_is_initial(d0[0])):
        s.set_order_parameter(c, x)
    for _ in range(2):
        s.compute("fp64")
    d = s.order_parameter(0)
    a = s.vector_potential(0)
    s.set_vector_potential(a[0], a[1])
    t = time.perf_counter()
    r0, g, ga = s.residual_gradient()
    t = time.perf_counter() - t
    dom = s.domain() > 0
    rng = np.random.default_rng(13)
    pts = [tuple(p) for p in np.argwhere(dom)[rng.choice(int(dom.sum()), 24, replace=False)]]
    n = s.N0
    pts_a = pts[:14] + [(1, 1), (n // 2, 2)]
    gmax, gamax = float(np.max(np.abs(g))), float(np.max(np.abs(ga)))
    out = {}
    for step in (1e-6, 1e-4):
        worst = 0.0
        for (y, x) in pts:
            for c in range(2):
                for unit in (1.0, 1j):
                    dd = d[c].copy()
                    dd[y, x] += step * unit
                    s.set_order_parameter(c, dd)
                    fp, _, _ = s.residual_gradient()
                    dd[y, x] -= 2 * step * unit
                    s.set_order_parameter(c, dd)
                    fm, _, _ = s.residual_gradient()
                    s.set_order_parameter(c, d[c])
                    an = g[c][y, x].real if unit == 1.0 else g[c][y, x].imag
                    worst = max(worst, abs((fp - fm) / (2 * step) - an))
        worst_a = 0.0
        for (y, x) in pts_a:
            for f in range(2):
                aa = a.copy()
                aa[f, y, x] += step
                s.set_vector_potential(aa[0], aa[1])
                fp, _, _ = s.residual_gradient()
                aa[f, y, x] -= 2 * step
                s.set_vector_potential(aa[0], aa[1])
                fm, _, _ = s.residual_gradient()
                worst_a = max(worst_a, abs((fp - fm) / (2 * step) - ga[f, y, x]))
        s.set_vector_potential(a[0], a[1])
        out[step] = (worst / gmax, worst_a / gamax)
    print("full-size d+is square: R %.6e, one residual gradient %.1f ms, |grad| max %.3e, in A "
          "%.3e; FD error of max (Delta, A): %.2e, %.2e at step 1e-6; %.2e, %.2e at step 1e-4" % (
              r0, 1e3 * t, gmax, gamax, out[1e-6][0], out[1e-6][1], out[1e-4][0], out[1e-4][1]))
    # measured 4e-5 / 2e-5 at step 1e-6 (rounding of R, 2.6e4 here, over the
    # step) and 5e-7 / 2e-7 at step 1e-4
    assert out[1e-6][0] < 2e-4 and out[1e-6][1] < 2e-4
    assert out[1e-4][0] < 2e-6 and out[1e-4][1] < 2e-6
    s.close()


@pytest.mark.parametrize("name", sorted(CASES))
def test_jacobian_vector_matches_oracle_differences(product, oracle, name):
    # sc_jacobian_vector (the tangent-linear sweep) against central
    # differences of the oracle's map G (one plain iteration, boundary
    # in-values fixed) along a seeded direction in Delta and A, step 1e-5:
    # the differences resolve J v to about 1e-9 of its largest entry in
    # Delta, 1e-8 in A (measured 2e-9 and 2e-8 to 6e-8 at step 1e-6)
    cfg = CASES[name]()
    d, a = state(product, cfg)
    s, o = fresh_pair(product, oracle, cfg, d, a)
    o.plain_map()  # the oracle's in-values from this state, as the product's
    rng = np.random.default_rng(21)
    n = s.N0
    vd = rng.standard_normal((s.Nc, n, n)) + 1j * rng.standard_normal((s.Nc, n, n))
    vd *= 0.1 * np.max(np.abs(d))
    va = rng.standard_normal((2, n, n)) * (0.1 * max(np.max(np.abs(a)), 1e-3)) if lam(cfg) > 0 \
        else None
    jd, ja = s.jacobian_vector(vd, va)
    eps = 1e-5
    outs = []
    for sign in (1.0, -1.0):
        for c in range(s.Nc):
            o.set_order_parameter(c, d[c] + sign * eps * vd[c])
        if va is not None:
            ap = a + sign * eps * va
            o.set_vector_potential(ap[0], ap[1])
        outs.append(o.plain_map())
    fd_d = (outs[0][0] - outs[1][0]) / (2 * eps)
    err = float(np.max(np.abs(jd - fd_d)) / np.max(np.abs(fd_d)))
    err_a = 0.0
    if va is not None:
        fd_a = (outs[0][1] - outs[1][1]) / (2 * eps)
        err_a = float(np.max(np.abs(ja - fd_a)) / np.max(np.abs(fd_a)))
    else:
        assert not np.any(ja)
    print("%s: J v against the oracle's differences: %.2e of max in Delta, %.2e in A" % (
        name, err, err_a))
    assert err < 1e-7 and err_a < 2e-7
    s.close()



@pytest.mark.parametrize("name", sorted(CASES))
def test_checkpointed_reverse_sweeps_give_the_same_gradients(product, name):
    # sc_set_adjoint_checkpoints: the coherence functions kept only every k
    # sites and each segment integrated again must give the gradients of
    # keeping every site (the same arithmetic; the transposes' atomic sums
    # add in another order)
    cfg = CASES[name]()
    d, a = state(product, cfg)
    s = product.Simulation(cfg)
    for c in range(d.shape[0]):
        s.set_order_parameter(c, d[c])
    if lam(cfg) > 0:
        s.set_vector_potential(a[0], a[1])
    s.set_adjoint_checkpoints(0)
    f0, g0, ga0 = s.free_energy_gradient(vector_potential=True)
    r0, h0, ha0 = s.residual_gradient()
    for k in (1, 3, 7, -1):
        s.set_adjoint_checkpoints(k)
        f, g, ga = s.free_energy_gradient(vector_potential=True)
        r, hh, ha = s.residual_gradient()
        errs = (abs(f - f0) / abs(f0), np.max(np.abs(g - g0)) / np.max(np.abs(g0)),
                abs(r - r0) / r0, np.max(np.abs(hh - h0)) / np.max(np.abs(h0)))
        print("%s, checkpoints every %d: differences %s" % (name, k, ["%.1e" % e for e in errs]))
        assert errs[0] <= 1e-14 and errs[2] <= 1e-14
        assert errs[1] <= 1e-13 and errs[3] <= 1e-13
        if lam(cfg) > 0:
            assert np.max(np.abs(ga - ga0)) <= 1e-13 * np.max(np.abs(ga0))
            assert np.max(np.abs(ha - ha0)) <= 1e-13 * np.max(np.abs(ha0))
    with pytest.raises(product.SupercongaError):
        s.set_adjoint_checkpoints(-2)
    s.close()
