"""The oracle's Keldysh spectroscopy (src/postprocess.cu) against what the
mathematics fixes.

* A spatially uniform s-wave gap in a grain with specular edges is a fixed
  point of every step: the bulk coherence functions at E + i delta are
  exact solutions along every trajectory and specular reflection maps them
  onto themselves (the gap has no momentum dependence).  The LDOS at every
  inside site is then the broadened BCS density of states
  Re[-i eps / sqrt(|Delta|^2 - eps^2)] = Re[eps / sqrt(eps^2 - |Delta|^2)],
  eps = E + i delta (LocalGreens.h:18-29 with RiccatiSolutions.h:36-54).
* Without a gap the coherence functions vanish and the LDOS is the angle
  sum of the integrand factors, 1 (the normal-state density of states).
* Energies never couple: each energy's trajectories, boundary condition and
  fixed point are its own, so the converged LDOS does not depend on how the
  energies are split into blocks (IntegrationIteratorKeldysh.h:152-180).
* The reference's own DOS test (tests/physics/dos_test.cu): s-wave disc and
  d-wave square at the bulk gap, grain-averaged LDOS equal to the angle
  average of the bulk density of states to 1e-8.
"""
import numpy as np
import pytest

import cases
from oracle import pyoracle as po

SPEC = {"energy_min": 0.0, "energy_max": 0.5, "num_energies": 9, "energy_broadening": 0.01}


def uniform_oracle(delta0, flux=0.0, criterion=1e-10, angles=16):
    cfg = cases.base(temperature=0.5, angles=angles, ppxi=3.0, criterion=criterion, flux=flux)
    cfg["numerics"]["norm"] = "l2"
    o = po.Oracle(cfg)
    n = o.N0
    o.set_order_parameter(0, np.full((n, n), delta0, dtype=complex))
    o.set_vector_potential(np.zeros((n, n)), np.zeros((n, n)))
    return o


def test_uniform_swave_gives_bcs_density_of_states():
    d0 = 0.2
    o = uniform_oracle(d0)
    o.ldos_setup(SPEC, 3)
    assert o.ldos_run(-1, 0, 50)
    ldos = o.ldos()
    inside = o.domain() > 0
    eps = np.linspace(SPEC["energy_min"], SPEC["energy_max"], SPEC["num_energies"]) \
        + 1j * SPEC["energy_broadening"]
    bcs = (-1j * eps / np.sqrt(d0 ** 2 - eps ** 2)).real
    for e in range(SPEC["num_energies"]):
        assert np.max(np.abs(ldos[e][inside] - bcs[e])) < 1e-11 * max(bcs[e], 1.0)
        assert np.all(ldos[e][~inside] == 0.0)
    # the gap edge and the in-gap suppression are there
    assert bcs[0] < 0.06 and bcs[4] > 1.5


def test_normal_state_density_of_states_is_one():
    o = uniform_oracle(0.0)
    o.ldos_setup(SPEC, -1)
    assert o.ldos_run(0, 0, 10)
    ldos = o.ldos()
    inside = o.domain() > 0
    assert np.max(np.abs(ldos[:, inside] - 1.0)) < 1e-13


def test_blocks_do_not_couple():
    cfg = cases.base(temperature=0.3, angles=16, ppxi=3.0, criterion=1e-9,
                     order_parameter={"dx2-y2": {"critical_temperature": 1.0}},
                     rotation=0.125, components=[cases.regular_polygon_unit(6)])
    cfg["numerics"]["norm"] = "l2"
    base = po.Oracle(cfg)
    base.initial_bulk()
    for _ in range(3):
        base.compute()
    delta = base.delta()[0]
    spec = dict(SPEC, num_energies=6, energy_broadening=0.05)
    out = []
    for per_block in (-1, 2, 3):
        o = po.Oracle(cfg)
        o.set_order_parameter(0, delta)
        o.set_vector_potential(np.zeros((o.N0, o.N0)), np.zeros((o.N0, o.N0)))
        o.ldos_setup(spec, per_block)
        assert o.ldos_run(0, 0, 400)
        out.append(o.ldos())
    scale = np.max(np.abs(out[0]))
    for other in out[1:]:
        assert np.max(np.abs(other - out[0])) < 1e-7 * scale
    # a pair-breaking edge: the LDOS is not the bulk one everywhere
    assert np.ptp(out[0][0][base.domain() > 0]) > 1e-2


def test_iteration_bookkeeping():
    # Context::isPartiallyConverged / isConverged and the block pointer
    o = uniform_oracle(0.2, criterion=1e-8)
    o.ldos_setup(SPEC, 4)
    assert o.get_int("ldos_block_start") == 0
    o.ldos_compute(burnin=False)
    # the first iteration compares against zeros: residual 1-ish, not converged
    assert o.get("residual_ldos") > 1e-3
    assert not o.get_int("ldos_partially_converged")
    for _ in range(5):
        o.ldos_compute(burnin=False)
        if o.get_int("ldos_block_start") > 0:
            break
    assert o.get_int("ldos_block_start") == 4
    assert not o.get_int("ldos_converged")
    assert o.ldos_run(0, 0, 20)
    assert o.get_int("ldos_converged")
    assert o.get_int("ldos_block_start") >= SPEC["num_energies"]


def test_ldos_setup_rejects_bad_energies():
    o = uniform_oracle(0.2)
    with pytest.raises(RuntimeError):
        o.ldos_setup(dict(SPEC, num_energies=0), -1)


# tests/physics/dos_test.cu:32-241: a grain of width 5 at 2 points per
# coherence length, 160 angles, 32 energies in [0, 0.5] with broadening
# 0.008, T = 0.5, L-infinity norm, at most 42 iterations, the order parameter
# at its bulk value; the grain-averaged LDOS must equal the angle average of
# the bulk density of states Im(eps / sqrt(|Delta eta|^2 - eps^2)) to 1e-8.
DOS_SPEC = {"energy_min": 0.0, "energy_max": 0.5, "num_energies": 32,
            "energy_broadening": 0.008}


def boole_weight(k, n):
    # FermiSurface::getAngleWeight (FermiSurface.h:119-138) for n % 4 == 0
    assert n % 4 == 0
    return (28.0 if k % 4 == 0 else 24.0 if k % 2 == 0 else 64.0) / 45.0


@pytest.mark.parametrize("symmetry", ["s", "dx2-y2"])
def test_reference_dos_known_answer(symmetry):
    n_angles = 160
    comps = [{"disc": {"radius": 0.5}}] if symmetry == "s" else [
        {"regular_polygon": {"num_edges": 4}}]
    cfg = cases.base(temperature=0.5, angles=n_angles, ppxi=2.0, criterion=1e-8,
                     norm="linf", width=5.0, components=comps,
                     order_parameter={symmetry: {"critical_temperature": 1.0}})
    o = po.Oracle(cfg)
    o.initial_bulk()
    o.set_vector_potential(np.zeros((o.N0, o.N0)), np.zeros((o.N0, o.N0)))
    o.ldos_setup(DOS_SPEC, -1)
    assert o.ldos_run(0, 0, 42)
    ldos = o.ldos()
    inside = o.domain() > 0
    d0 = o.get("delta_bulk_0")
    eps = np.linspace(0.0, 0.5, 32) + 0.008j
    phi = np.pi * 2 * np.arange(n_angles) / n_angles + 0.5 * np.pi
    if symmetry == "s":
        eta = np.ones(n_angles)
    else:
        raw = np.cos(phi) ** 2 - np.sin(phi) ** 2
        w = np.array([boole_weight(k, n_angles) for k in range(n_angles)]) / n_angles
        eta = raw / np.sqrt(np.sum(w * raw * raw))  # BasisFunction::normalize
    w = np.array([boole_weight(k, n_angles) for k in range(n_angles)]) / n_angles
    for e in range(32):
        omega = np.sqrt(np.abs(d0 * eta) ** 2 - eps[e] ** 2)
        expected = float(np.sum(w * (eps[e] / omega).imag))
        got = float(np.sum(ldos[e][inside]) / np.count_nonzero(inside))
        assert abs(got - expected) <= 1e-8 * abs(expected), (e, got, expected)
