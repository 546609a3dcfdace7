"""The reference's own physics known-answer tests, run on the product in
FP64 with the reference's parameters, iteration limits and tolerances:

* tests/physics/swave_disc_test.cu: an s-wave disc (R = 20 xi0) in 5 flux
  quanta with penetration depth 5 xi0 converges within 100
  Barzilai-Borwein iterations, and its free energy per area is the London
  result F_bulk + F_kinetic + F_magnetic (modified Bessel functions of R/k,
  k = lambda / sqrt(Yoshida)) to 1e-3;
* tests/physics/swave_annulus_solenoid_gauge_test.cu: an s-wave annulus
  (R = 20, r = 10 xi0) with phase winding 1 in the solenoid gauge converges
  at flux -1.5, -1, -0.5, and its free energy per area is
  F_bulk + Y ln(R/r) / (4 (R^2 - r^2)) (Phi - q w)^2 to 1e-3, lowest at
  Phi = q w;
* tests/physics/current_phase_gradient_test.cu: after one iteration of a
  disc with phase winding 1 and no field, j_i d_i phase has the sign of
  q w everywhere (zero allowed);
* tests/physics/penetration_depth_test.cu: the London fit of the flux
  through a disc in a weak field gives the penetration depth to 10%;
* tests/physics/low_temp_free_energy_test.cu: at T = 0.001 (281 Ozaki
  poles) an s-wave disc, an s-wave square and a d-wave square keep the
  T = 0 bulk gap on average to 1e-5, carry no current (below machine
  epsilon), and have F / area = -Delta_bulk^2 / 2 to 1e-4;
* tests/physics/swave_annulus_radial_gauge_test.cu: an annulus (R = 20,
  r = 10, lambda = 30) with phase winding -2 in one flux quantum has the
  London free energy (1e-3), magnetic free energy (2e-2) and magnetic
  moment (1e-3), the hole's field energy written as that of its uniform
  field (tests/test_oracle_reference_physics.py:annulus_radial_expected);
* tests/physics/current_linear_response_test.cu: a square with a uniform
  vector potential A = (0.01, 0.01) carries j = -Y(T) A at its centre after
  one iteration, to 1e-2, at T = 0.1 and 0.9.

The bulk gap, bulk free energy and Yoshida function come from the oracle,
whose values tests/test_oracle_closed_forms.py pins on the reference's
tests/order_parameter_bulk_test.cu.
"""
import math

import numpy as np
import pytest
from scipy import special

import cases

pytestmark = pytest.mark.gpu


def bulk_terms(oracle, temperature, cutoff):
    d = oracle.bulk_gap(temperature)
    return oracle.bulk_free_energy(temperature, d), oracle.yoshida(temperature, d, cutoff)


def test_reference_swave_disc_free_energy(product, oracle):
    # tests/physics/swave_disc_test.cu:60-160
    T, cutoff, lam, flux, R = 0.2, 16.0, 5.0, 5.0, 20.0
    cfg = cases.base(temperature=T, flux=flux, lam=lam, angles=16, ppxi=2.7, cutoff=cutoff,
                     norm="linf", criterion=1e-3, accelerator="barzilai-borwein",
                     width=2 * R)
    s = product.Simulation(cfg)
    converged = s.run(0, 0, 100, "fp64")
    f = s.results()["free_energy_per_area"]
    # the analytic London result (helper::computeFreeEnergyAnalytic)
    f_bulk, y = bulk_terms(oracle, T, cutoff)
    k = lam / math.sqrt(y)
    area = math.pi * R * R
    b = math.pi * flux / area
    factor = 2 * math.pi * y * (k * b) ** 2 / (2 * area)
    i0, i1 = special.iv(0, R / k), special.iv(1, R / k)
    tmp = R * i1 / i0
    f_mag = factor * (R * R - 2 * k * tmp - 0.5 * tmp * tmp)
    f_kin = factor * 0.5 * R * (-R * i0 * i0 + 2 * k * i0 * i1 + R * i1 * i1) / (i0 * i0)
    expected = f_bulk + f_kin + f_mag
    print("s-wave disc: %d iterations, F/area %.6f, London %.6f, rel %.2e" % (
        s.results()["iterations"], f, expected, abs(f - expected) / abs(expected)))
    assert converged
    assert abs(f - expected) <= 1e-3 * abs(expected)
    s.close()


def test_reference_swave_annulus_solenoid_free_energy(product, oracle):
    # tests/physics/swave_annulus_solenoid_gauge_test.cu:38-170
    T, cutoff, R, r = 0.2, 16.0, 20.0, 10.0
    winding, charge = 1.0, -1
    qw = charge * winding
    fluxes = [qw - 0.5, qw, qw + 0.5]
    f_bulk, y = bulk_terms(oracle, T, cutoff)
    factor = 0.25 * y * math.log(R / r) / (R * R - r * r)
    energies = []
    for flux in fluxes:
        cfg = cases.base(
            order_parameter={"s": {"critical_temperature": 1.0,
                                   "vortices": [{"center_x": 0.0, "center_y": 0.0,
                                                 "winding_number": winding}]}},
            temperature=T, flux=flux, lam=-1.0, angles=16, ppxi=3.5, cutoff=cutoff,
            norm="linf", criterion=1e-4, accelerator="barzilai-borwein", step_size=1.0,
            gauge="solenoid", width=2 * R,
            components=[{"disc": {"radius": 0.5}},
                        {"disc": {"radius": 0.5 * r / R, "add": False}}])
        s = product.Simulation(cfg)
        converged = s.run(0, 0, 100, "fp64")
        f = s.results()["free_energy_per_area"]
        expected = f_bulk + factor * (flux - qw) ** 2
        print("annulus, flux %+.1f: %d iterations, F/area %.6f, expected %.6f, rel %.2e" % (
            flux, s.results()["iterations"], f, expected, abs(f - expected) / abs(expected)))
        assert converged
        assert abs(f - expected) <= 1e-3 * abs(expected)
        energies.append(f)
        s.close()
    assert energies[1] < energies[0] and energies[1] < energies[2]


def phase_gradient(delta, h=1.0):
    """GridFunctions::complexPhaseGradient (GridFunctions_CPU.h:343-410):
    central differences of the phase from Re and Im, zero on the border and
    where 2 h |delta|^2 <= 1e-5."""
    re, im = delta.real, delta.imag
    gx = np.zeros_like(re)
    gy = np.zeros_like(re)
    denom = 2.0 * h * (re * re + im * im)
    c = (slice(1, -1), slice(1, -1))
    ok = denom[c] > 1e-5
    inv = np.where(ok, 1.0 / np.where(ok, denom[c], 1.0), 0.0)
    dx_im = im[1:-1, 2:] - im[1:-1, :-2]
    dx_re = re[1:-1, 2:] - re[1:-1, :-2]
    dy_im = im[2:, 1:-1] - im[:-2, 1:-1]
    dy_re = re[2:, 1:-1] - re[:-2, 1:-1]
    gx[c] = (dx_im * re[c] - im[c] * dx_re) * inv
    gy[c] = (dy_im * re[c] - im[c] * dy_re) * inv
    return gx, gy


def phase_gradient_case():
    # tests/physics/current_phase_gradient_test.cu:33-75
    return cases.base(order_parameter={"s": {"critical_temperature": 1.0,
                                             "vortices": [{"center_x": 0.0, "center_y": 0.0,
                                                           "winding_number": 1.0}]}},
                      temperature=0.9, flux=0.0, lam=-1.0, angles=16, ppxi=2.0, cutoff=16.0,
                      norm="linf", criterion=1e-5, accelerator="barzilai-borwein", width=10.0)


def check_overlap_sign(j, delta, expected_sign):
    gx, gy = phase_gradient(delta)
    for comp, g in ((j[0], gx), (j[1], gy)):
        overlap = comp * g
        # the reference's check: neither extreme has the wrong sign
        assert np.sign(overlap.min()) != -expected_sign
        assert np.sign(overlap.max()) != -expected_sign
        # and the current does flow along the phase gradient
        assert np.max(np.abs(overlap)) > 0.0


def test_reference_current_follows_phase_gradient(product):
    s = product.Simulation(phase_gradient_case())
    s.compute("fp64")
    check_overlap_sign(s.current(), s.order_parameter(0)[0], expected_sign=-1)
    s.close()


def test_reference_penetration_depth(product):
    # tests/physics/penetration_depth_test.cu: the flux through a disc in a
    # weak field fitted with the London expression gives the penetration
    # depth back to 10% (the reference's tolerance); the oracle's fit is
    # 1.107 (tests/test_oracle_reference_physics.py)
    from test_oracle_reference_physics import penetration_case, penetration_depth_from
    s = product.Simulation(penetration_case())
    converged = s.run(0, 0, 100, "fp64")
    d = s.dims
    k = penetration_depth_from(s.vector_potential(0), s.domain(), d.h, d.num_grain_sites)
    print("penetration depth fit %.4f after %d iterations" % (k, s.results()["iterations"]))
    assert converged
    assert abs(k - 1.0) <= 0.1 * max(k, 1.0)
    s.close()


@pytest.mark.parametrize("symmetry,shape", [("s", "disc"), ("s", "square"),
                                            ("dx2-y2", "square")])
def test_reference_low_temperature_bulk_values(product, symmetry, shape):
    from test_oracle_reference_physics import check_low_temperature, low_temperature_case
    s = product.Simulation(low_temperature_case(symmetry, shape))
    converged = s.run(0, 1, 50, "fp64")
    r = s.results()
    print("%s %s: %d iterations" % (symmetry, shape, r["iterations"]))
    check_low_temperature(symmetry, s.order_parameter(0)[0], s.domain(), s.current(),
                          s.dims.num_grain_sites, r["free_energy_per_area"], converged)
    s.close()


def test_reference_swave_annulus_radial_free_energy_and_moment(product):
    from test_oracle_reference_physics import (annulus_radial_case, check_annulus_radial,
                                               magnetic_moment_per_area)
    s = product.Simulation(annulus_radial_case())
    converged = s.run(0, 0, 100, "fp64")
    r = s.results()
    d = s.dims
    print("annulus (radial): %d iterations" % r["iterations"])
    check_annulus_radial(r["free_energy_per_area"], r["magnetic_free_energy_per_area"],
                         magnetic_moment_per_area(s.current(), d.h, d.num_grain_sites),
                         converged)
    s.close()


@pytest.mark.parametrize("temperature", [0.1, 0.9])
def test_reference_current_linear_response(product, oracle, temperature):
    from test_oracle_reference_physics import (LINEAR_RESPONSE_A, check_linear_response,
                                               linear_response_case)
    s = product.Simulation(linear_response_case(temperature))
    n = s.N0
    a = np.full((n, n), LINEAR_RESPONSE_A)
    s.set_vector_potential(a, a)
    s.compute("fp64")
    d = oracle.bulk_gap(temperature)
    check_linear_response(s.current(), temperature, oracle.yoshida(temperature, d, 16.0))
    s.close()
