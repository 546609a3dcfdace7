"""The stages after the sweep replayed from a captured hipGraph
(sc_set_graphs, on by default) against the same stages launched eagerly:
the same kernels in the same order, so every field and every scalar of
every iteration must agree bit for bit.  Several iterations run between the
replays, each with the eager accelerator step and the read-back of its
residuals, so a graph node that went stale after its first replay would
show."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

CASES = {
    "disc_field_congacc": lambda: cases.accelerator_case("congacc"),
    "disc_field_anderson": lambda: cases.accelerator_case("anderson"),
    "d_plus_s_hole": cases.mixed_small,
    "screened_picard": lambda: dict(cases.accelerator_case("picard"),
                                    accelerator={"name": "picard", "step_size": 1.0,
                                                 "vector_potential_screening": 1.0}),
}


def snapshot(s):
    r = s.results()
    fields = [s.order_parameter(0), s.order_parameter(1), s.current(),
              s.free_energy_density(), s.vector_potential(0), s.vector_potential(1)]
    scal = {k: r[k] for k in r if k.startswith("residual") or k.startswith("free_energy")
            or k.startswith("magnetic")}
    return fields, scal


def assert_same(a, b, what):
    fa, sa = a
    fb, sb = b
    for i, (x, y) in enumerate(zip(fa, fb)):
        assert np.array_equal(x, y), (what, i, float(np.max(np.abs(x - y))))
    for k in sa:
        assert sa[k] == sb[k] or (np.isnan(sa[k]) and np.isnan(sb[k])), (what, k, sa[k], sb[k])


@pytest.mark.parametrize("name", sorted(CASES))
def test_graph_replay_equals_eager(product, name):
    cfg = CASES[name]()
    graph, eager = product.Simulation(cfg), product.Simulation(cfg)
    eager.set_graphs(False)
    d0 = graph.order_parameter(0)
    for c, d in enumerate(d0):
        eager.set_order_parameter(c, d)
    for it in range(8):
        graph.compute("fp64")
        eager.compute("fp64")
        assert_same(snapshot(graph), snapshot(eager), "%s it %d" % (name, it))
    # a burn-in and mixed-precision iterations in between, then FP64 again
    graph.burnin("fp64")
    eager.burnin("fp64")
    for it in range(3):
        graph.compute("mixed")
        eager.compute("mixed")
        assert_same(snapshot(graph), snapshot(eager), "%s mixed it %d" % (name, it))
    graph.compute("fp64")
    eager.compute("fp64")
    assert_same(snapshot(graph), snapshot(eager), "%s after" % name)
    # switching the graphs off mid-run keeps the state
    graph.set_graphs(False)
    graph.compute("fp64")
    eager.compute("fp64")
    assert_same(snapshot(graph), snapshot(eager), "%s graphs off" % name)
    graph.close()
    eager.close()


def test_graph_replay_with_angle_hooks(product):
    # the angle-by-angle iteration finishes with its own gather: a graph of
    # the fused iteration must not be replayed for it, and the reverse
    cfg = cases.accelerator_case("congacc")
    graph, eager = product.Simulation(cfg), product.Simulation(cfg)
    eager.set_graphs(False)
    for c, d in enumerate(graph.order_parameter(0)):
        eager.set_order_parameter(c, d)
    na = graph.dims.num_angles
    for it in range(6):
        for s in (graph, eager):
            if it % 2:
                s.compute("fp64")
            else:
                s.integration_begin()
                for k0 in range(0, na, 5):
                    s.sweep_angles(k0, min(k0 + 5, na))
                    s.integrate_angles(k0, min(k0 + 5, na))
                s.integration_end()
        assert_same(snapshot(graph), snapshot(eager), "hooks it %d" % it)


def test_batch_graph_replay_equals_eager(product):
    cfgs = [cases.accelerator_case("congacc"), cases.mixed_small(),
            cases.base(temperature=0.4, flux=0.5, lam=6.0, angles=16, ppxi=3.0, width=8.0)]
    graph, eager = product.Batch(cfgs), product.Batch(cfgs)
    eager.set_graphs(False)
    for i in range(len(cfgs)):
        for c, d in enumerate(graph.grain(i).order_parameter(0)):
            eager.grain(i).set_order_parameter(c, d)
    for it in range(6):
        graph.compute("fp64")
        eager.compute("fp64")
        for i in range(len(cfgs)):
            assert_same(snapshot(graph.grain(i)), snapshot(eager.grain(i)),
                        "batch grain %d it %d" % (i, it))
