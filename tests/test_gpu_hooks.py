"""The reference's module hooks through the C ABI: the iteration driven
angle by angle (IntegrationIteratorOzaki::compute calling the RiccatiSolver
and the ComputeProperty hooks) must give what one fused sc_compute gives,
bit for bit, and the accelerator step on caller-supplied fields
(Accelerator<T>::compute) what the context's own step gives."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def twins(product, cfg):
    a = product.Simulation(cfg)
    b = product.Simulation(cfg)
    for c, d in enumerate(a.order_parameter(0)):
        b.set_order_parameter(c, d)
    return a, b


def same_state(a, b):
    assert np.array_equal(a.order_parameter(0), b.order_parameter(0))
    assert np.array_equal(a.order_parameter(1), b.order_parameter(1))
    assert np.array_equal(a.current(), b.current())
    assert np.array_equal(a.vector_potential(1), b.vector_potential(1))
    assert np.array_equal(a.free_energy_density(), b.free_energy_density())
    for io in (0, 1):
        for d in (0, 1):
            assert np.array_equal(a.boundary(io, d), b.boundary(io, d))
    ra, rb = a.results(), b.results()
    for k in ("free_energy", "residual", "residual_boundary", "residual_order_parameter",
              "residual_current", "residual_vector_potential", "iterations"):
        assert ra[k] == rb[k], k


@pytest.mark.parametrize("chunk", [1, 5])
def test_angle_by_angle_equals_fused_iteration(product, chunk):
    cfg = cases.mixed_small()  # two components, field, induced potential, hole
    fused, hooks = twins(product, cfg)
    na = fused.Na
    # a burn-in and three iterations; the angle ranges swept out of order
    # (the sweep of an angle reads only the previous iteration's in-values)
    # and integrated in order
    fused.burnin("fp64")
    hooks.integration_begin(burnin=True)
    for k0 in range(0, na, chunk):
        hooks.sweep_angles(k0, min(k0 + chunk, na), burnin=True)
    hooks.integration_end(burnin=True)
    same_state(fused, hooks)
    for it in range(3):
        fused.compute("fp64")
        hooks.integration_begin()
        starts = list(range(0, na, chunk))
        for k0 in reversed(starts):
            hooks.sweep_angles(k0, min(k0 + chunk, na))
        for k0 in starts:
            hooks.integrate_angles(k0, min(k0 + chunk, na))
        hooks.integration_end()
        print("iteration %d: residual %.3e" % (it, hooks.results()["residual"]), flush=True)
        same_state(fused, hooks)


def test_angle_hooks_refuse_out_of_order_use(product):
    s = product.Simulation(cases.accelerator_case("picard"))
    with pytest.raises(product.SupercongaError):
        s.sweep_angles(0, 1)  # outside begin/end
    s.integration_begin()
    s.sweep_angles(0, s.Na)
    with pytest.raises(product.SupercongaError):
        s.integrate_angles(1, 2)  # angle 0 first
    with pytest.raises(product.SupercongaError):
        s.integration_end()  # not every angle integrated
    s.integrate_angles(0, s.Na)
    s.integration_end()
    assert s.results()["iterations"] == 1


@pytest.mark.parametrize("accelerator", ["anderson", "congacc", "barzilai-borwein",
                                         "polyak", "broyden"])
def test_accelerator_step_on_fields_equals_context_step(product, accelerator):
    # Context<T>::compute applies the accelerator to (iterate, map result);
    # the same pairs handed to sc_accelerator_step of a second context with
    # the same config give the same accelerated iterates
    cfg = cases.accelerator_case(accelerator)
    run, acc = twins(product, cfg)
    nf = acc.num_fields()
    assert nf == 2 * run.Nc + 2
    lam2 = cfg["physics"]["penetration_depth"] ** 2

    def stack(sim, which):
        d = sim.order_parameter(which)
        a = sim.vector_potential(which) * lam2
        planes = []
        for c in range(sim.Nc):
            planes += [d[c].real, d[c].imag]
        return np.stack(planes + [a[0], a[1]])

    run.compute("fp64")
    for it in range(4):
        x, g = stack(run, 0), stack(run, 1)
        out = acc.accelerator_step(x, g)
        run.compute("fp64")  # applies the step first
        nxt = stack(run, 0)
        err = float(np.max(np.abs(out - nxt)) / np.max(np.abs(nxt)))
        print("%s step %d: difference %.2e" % (accelerator, it, err), flush=True)
        assert err <= 1e-13
