"""Parity of the HIP path with the oracle at the benchmark grains' full
sizes (superconga_amd.workloads), through the C ABI.

Both sides start from the product's own seeded initial guess (bulk gap with
1% noise, the vortex phase winding where the config has one), handed to the
oracle as arrays, the way the reference hands a state over through its
load_path CSV files.  The bar is the FP64 one of test_gpu_parity.py: every
grid point of order parameter, current, vector potential and free-energy
density within 1e-10 of the oracle relative to the field's largest
magnitude, the free energy within 1e-10 relative, residuals to 1e-10 with
their absolute floors.  FP32 stepping states its error against the FP64
oracle per field.

The oracle takes about a minute per octagon iteration on 16 host threads
(5.1e9 trajectory steps plus the direct O(N0^4) vector-potential
convolution), so the octagon and the d + is square are compared over one
iteration; the s-wave disc is also run to its criterion on both sides.
"""
import numpy as np
import pytest

from superconga_amd import workloads
from test_gpu_parity import TOL, compare_state, rel, srel

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-4


def start(product, oracle, cfg, initial=None):
    """A product grain and an oracle context on the product's initial guess
    (optionally transformed by `initial`)."""
    s = product.Simulation(cfg)
    d0 = list(s.order_parameter(0))
    if initial is not None:
        d0 = initial(d0)
    o = oracle.Oracle(cfg)
    for c, d in enumerate(d0):
        s.set_order_parameter(c, d)
        o.set_order_parameter(c, d)
    return s, o, d0


def expect_dims(s, o, n0, n1, ne, na):
    assert (s.N0, s.N1, s.Ne, s.Na) == (n0, n1, ne, na)
    assert (o.N0, o.N1, o.Ne, o.Na) == (n0, n1, ne, na)


def test_swave_disc_iterations_and_run(product, oracle):
    # s-wave disc R = 10 xi0, 10 points per xi0, 32 angles, T = 0.5, Picard
    cfg = workloads.swave_disc()
    s, o, _ = start(product, oracle, cfg)
    expect_dims(s, o, 206, 292, 9, 32)
    for it in range(3):
        print("swave_disc iteration %d" % it, flush=True)
        s.compute("fp64")
        o.compute()
        compare_state(s, o, cfg, what="swave_disc it %d" % it)
    # then both run on to the L2 criterion 1e-6 and must stop together
    conv = s.run(0, 0, 500, "fp64")
    oconv = o.run(0, 0, 500)
    r = s.results()
    print("swave_disc run: converged %s/%s after %d/%d iterations, residual %.3e/%.3e"
          % (conv, oconv, r["iterations"], o.get_int("num_iterations"), r["residual"],
             o.get("residual")), flush=True)
    assert conv and oconv
    assert r["iterations"] == o.get_int("num_iterations")
    assert rel(s.order_parameter(0), o.delta(False)) <= TOL
    assert srel(r["free_energy_per_area"], o.get("free_energy_per_area")) <= TOL


def test_vortex_disc_fp64_and_fp32_stepping(product, oracle):
    # s-wave disc R = 12.5 xi0, 20 points per xi0, 64 angles, T = 0.2,
    # 2 flux quanta, penetration depth 3, one vortex; exact vector-potential
    # kernel (vector_potential_error 0)
    cfg = workloads.vortex_disc()
    s, o, d0 = start(product, oracle, cfg)
    expect_dims(s, o, 506, 716, 14, 64)
    m = product.Simulation(cfg)
    m.set_order_parameter(0, d0[0])
    for it in range(2):
        print("vortex_disc iteration %d" % it, flush=True)
        s.compute("fp64")
        m.compute("mixed")
        o.compute()
        compare_state(s, o, cfg, what="vortex_disc it %d" % it)
        lam2 = cfg["physics"]["penetration_depth"] ** 2
        err = {
            "delta": rel(m.order_parameter(1), o.delta(True)),
            "current": rel(m.current(), o.current()),
            "vector_potential": rel(m.vector_potential(1), o.vector_potential(True) / lam2),
            "free_energy": srel(m.results()["free_energy"], o.get("free_energy")),
        }
        print("vortex_disc it %d, FP32 stepping + FP64 sums vs FP64 oracle: %s"
              % (it, {k: "%.2e" % v for k, v in err.items()}), flush=True)
        assert all(v < MIXED_TOL for v in err.values()), err


def test_octagon_one_iteration(product, oracle):
    # regular octagon, circumradius 25 xi0, 20 points per xi0, d_{x2-y2} at
    # 45 degrees, 128 angles, T = 0.1, cutoff 32, penetration depth 5
    cfg = workloads.octagon_dwave()
    s, o, _ = start(product, oracle, cfg)
    expect_dims(s, o, 930, 1316, 28, 128)
    print("octagon iteration 0 (oracle: about a minute)", flush=True)
    s.compute("fp64")
    o.compute()
    compare_state(s, o, cfg, what="octagon it 0")


def test_dplus_is_square_one_iteration(product, oracle):
    # square of side 30 xi0, 20 points per xi0, d_{x2-y2} + i 0.2 s, 128
    # angles, T = 0.3, one flux quantum, penetration depth 3
    cfg = workloads.dplus_is_square()
    s, o, _ = start(product, oracle, cfg,
                    initial=lambda d: workloads.d_plus_is_initial(d[0]))
    expect_dims(s, o, 607, 859, 12, 128)
    print("d+is square iteration 0", flush=True)
    s.compute("fp64")
    o.compute()
    compare_state(s, o, cfg, what="d+is square it 0")


def test_parameter_sweep_batch(product, oracle):
    # the 48 grains of the parameter sweep in one batch: every grain equal
    # to the same grain iterated alone, and the three smallest equal to the
    # oracle
    cfgs = workloads.parameter_sweep()
    b = product.Batch(cfgs)
    print("parameter sweep: %d grains" % len(b), flush=True)
    sizes = []
    starts = []
    for i in range(len(b)):
        d = b.grain(i).order_parameter(0)
        starts.append(d)
        sizes.append(b.grain(i).dims.traj_steps)
    smallest = sorted(range(len(b)), key=lambda i: sizes[i])[:3]
    oracles = {}
    for i in smallest:
        o = oracle.Oracle(cfgs[i])
        for c, d in enumerate(starts[i]):
            o.set_order_parameter(c, d)
        oracles[i] = o
    for it in range(2):
        b.compute("fp64")
        for i, o in oracles.items():
            o.compute()
            compare_state(b.grain(i), o, cfgs[i], what="sweep grain %d it %d" % (i, it))
    # each grain alone, from the same start, two iterations
    worst = 0.0
    for i, cfg in enumerate(cfgs):
        s = product.Simulation(cfg)
        for c, d in enumerate(starts[i]):
            s.set_order_parameter(c, d)
        for _ in range(2):
            s.compute("fp64")
        g = b.grain(i)
        e = max(rel(g.order_parameter(1), s.order_parameter(1)), rel(g.current(), s.current()),
                srel(g.results()["free_energy"], s.results()["free_energy"]))
        worst = max(worst, e)
        s.close()
    print("parameter sweep: largest difference batch vs single grain %.2e" % worst,
          flush=True)
    # the batch and the single grain may take different sweep shapes, whose
    # energy sums round differently
    assert worst <= 1e-12
