"""The oracle against the reference's own physics known-answer tests, with
the reference's parameters, iteration limits and tolerances (the same
checks run on the product in tests/test_gpu_reference_physics.py):

* tests/physics/swave_disc_test.cu: London free energy of an s-wave disc in
  5 flux quanta, 1e-3;
* tests/physics/swave_annulus_solenoid_gauge_test.cu: free energy of an
  annulus with phase winding 1, parabolic in the flux with its minimum at
  Phi = q w, 1e-3.
"""
import math

import numpy as np
from scipy import special

import cases
from oracle import pyoracle as po


def bulk_terms(temperature, cutoff):
    d = po.bulk_gap(temperature)
    return po.bulk_free_energy(temperature, d), po.yoshida(temperature, d, cutoff)


def test_swave_disc_london_free_energy():
    T, cutoff, lam, flux, R = 0.2, 16.0, 5.0, 5.0, 20.0
    cfg = cases.base(temperature=T, flux=flux, lam=lam, angles=16, ppxi=2.7, cutoff=cutoff,
                     norm="linf", criterion=1e-3, accelerator="barzilai-borwein",
                     width=2 * R)
    o = po.Oracle(cfg)
    o.initial_bulk()
    assert o.run(0, 0, 100)
    f = o.get("free_energy_per_area")
    f_bulk, y = bulk_terms(T, cutoff)
    k = lam / math.sqrt(y)
    area = math.pi * R * R
    b = math.pi * flux / area
    factor = 2 * math.pi * y * (k * b) ** 2 / (2 * area)
    i0, i1 = special.iv(0, R / k), special.iv(1, R / k)
    tmp = R * i1 / i0
    f_mag = factor * (R * R - 2 * k * tmp - 0.5 * tmp * tmp)
    f_kin = factor * 0.5 * R * (-R * i0 * i0 + 2 * k * i0 * i1 + R * i1 * i1) / (i0 * i0)
    expected = f_bulk + f_kin + f_mag
    assert abs(f - expected) <= 1e-3 * abs(expected)
    # the field and the screening matter: the bulk value alone is off
    assert abs(f - f_bulk) > 1e-3 * abs(f_bulk)


def test_swave_annulus_solenoid_parabola():
    T, cutoff, R, r = 0.2, 16.0, 20.0, 10.0
    qw = -1.0  # charge sign -1, phase winding 1
    f_bulk, y = bulk_terms(T, cutoff)
    factor = 0.25 * y * math.log(R / r) / (R * R - r * r)
    energies = []
    for flux in (qw - 0.5, qw, qw + 0.5):
        cfg = cases.base(temperature=T, flux=flux, lam=-1.0, angles=16, ppxi=3.5,
                         cutoff=cutoff, norm="linf", criterion=1e-4,
                         accelerator="barzilai-borwein", step_size=1.0, gauge="solenoid",
                         width=2 * R,
                         components=[{"disc": {"radius": 0.5}},
                                     {"disc": {"radius": 0.5 * r / R, "add": False}}])
        o = po.Oracle(cfg)
        o.initial_bulk()
        # Initial<T>::phaseWinding: the bulk gap with the phase of a winding
        # about the grain's centre
        n = o.N0
        c = (n - 1) / 2.0
        yy, xx = np.mgrid[0:n, 0:n]
        o.set_order_parameter(0, o.delta()[0] * np.exp(1j * np.arctan2(yy - c, xx - c)))
        assert o.run(0, 0, 100)
        f = o.get("free_energy_per_area")
        assert abs(f - (f_bulk + factor * (flux - qw) ** 2)) <= 1e-3 * abs(f)
        energies.append(f)
    assert energies[1] < energies[0] and energies[1] < energies[2]
