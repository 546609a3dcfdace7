"""Named benchmark grains in the simulation_config.json schema.

Each function returns a config dict that both the HIP path (Simulation) and
the reference's simulation binary (src/simulation.cu:64-463) accept; sizes
follow Parameters::setLatticeResolution (Parameters.h:422-458).
"""
import math


def _numerics(angles, ppxi, cutoff, criterion, norm="l2"):
    return {"convergence_criterion": criterion, "energy_cutoff": cutoff,
            "norm": norm, "num_fermi_momenta": angles,
            "points_per_coherence_length": ppxi, "vector_potential_error": 0.0,
            "num_iterations_burnin": 0, "num_iterations_min": 0,
            "num_iterations_max": 10000}


def octagon_dwave(penetration_depth=5.0, seed=1234, noise=0.01):
    """d-wave octagon: regular 8-gon of circumradius 25 xi0 (side
    2 R sin(pi/8) = 19.134 xi0), 20 points per xi0 (N0 = 930, N1 = 1316),
    d_{x2-y2} with the crystal axes at 45 degrees so that the straight edges
    are pair-breaking, 128 Fermi-surface angles, T = 0.1 Tc, cutoff 32
    (28 Ozaki poles), zero external flux with a self-consistent vector
    potential (penetration depth 5 xi0), initial Delta = bulk d-wave
    amplitude with 1% seeded noise, Anderson mixing with history 8."""
    side = 2.0 * 25.0 * math.sin(math.pi / 8.0)
    return {
        "physics": {"temperature": 0.1, "external_flux_quanta": 0.0,
                    "penetration_depth": penetration_depth, "gauge": "symmetric",
                    "charge_sign": -1, "crystal_axes_rotation": 0.125,
                    "order_parameter": {"dx2-y2": {
                        "critical_temperature": 1.0,
                        "initial_noise_stddev": noise}}},
        "geometry": [{"regular_polygon": {"num_edges": 8, "side_length": side,
                                          "rotation": 0.0, "center_x": 0.0,
                                          "center_y": 0.0, "add": True}}],
        "numerics": _numerics(128, 20.0, 32.0, 1e-8),
        "accelerator": {"name": "anderson", "rank": 8, "step_size": 1.0},
        "misc": {"seed": seed},
    }


def vortex_disc(seed=1234, noise=0.01):
    """s-wave disc of radius 12.5 xi0 at 20 points per xi0 (N0 = 506),
    64 angles, T = 0.2 Tc, cutoff 16, 2 flux quanta, penetration depth
    3 xi0, one vortex at the centre with 1% seeded noise."""
    return {
        "physics": {"temperature": 0.2, "external_flux_quanta": 2.0,
                    "penetration_depth": 3.0, "gauge": "symmetric",
                    "charge_sign": -1,
                    "order_parameter": {"s": {
                        "critical_temperature": 1.0, "initial_noise_stddev": noise,
                        "vortices": [{"center_x": 0.0, "center_y": 0.0,
                                      "winding_number": 1}]}}},
        "geometry": [{"disc": {"radius": 12.5, "center_x": 0.0, "center_y": 0.0,
                               "add": True}}],
        "numerics": _numerics(64, 20.0, 16.0, 1e-8),
        "accelerator": {"name": "congacc"},
        "misc": {"seed": seed},
    }


def swave_disc(seed=1234, noise=0.01):
    """s-wave disc of radius 10 xi0 at 10 points per xi0 (N0 = 206), 32
    angles, T = 0.5 Tc, cutoff 16, no field, Picard, L2 criterion 1e-6."""
    return {
        "physics": {"temperature": 0.5, "external_flux_quanta": 0.0,
                    "penetration_depth": -1.0, "gauge": "symmetric",
                    "charge_sign": -1,
                    "order_parameter": {"s": {
                        "critical_temperature": 1.0, "initial_noise_stddev": noise}}},
        "geometry": [{"disc": {"radius": 10.0, "center_x": 0.0, "center_y": 0.0,
                               "add": True}}],
        "numerics": _numerics(32, 10.0, 16.0, 1e-6),
        "accelerator": {"name": "picard", "step_size": 1.0},
        "misc": {"seed": seed},
    }


def dplus_is_square(seed=1234, noise=0.01, penetration_depth=3.0, s_critical_temperature=0.5):
    """Two-component d + is square: side 30 xi0 at 20 points per xi0
    (N0 = 607, N1 = 859), 128 angles, T = 0.3 Tc, cutoff 16 (12 Ozaki
    poles), one flux quantum, penetration depth 3 xi0, d_{x2-y2} (Tc 1) and
    s (Tc 0.5, a quarter turn out of phase).  The starting state puts the s
    amplitude at 0.2 of the d amplitude (d_plus_is_initial)."""
    return {
        "physics": {"temperature": 0.3, "external_flux_quanta": 1.0,
                    "penetration_depth": penetration_depth, "gauge": "symmetric",
                    "charge_sign": -1,
                    "order_parameter": {
                        "dx2-y2": {"critical_temperature": 1.0,
                                   "initial_noise_stddev": noise},
                        "s": {"critical_temperature": s_critical_temperature,
                              "initial_phase_shift": 0.25,
                              "initial_noise_stddev": noise}}},
        "geometry": [{"regular_polygon": {"num_edges": 4, "side_length": 30.0,
                                          "rotation": 0.0, "center_x": 0.0,
                                          "center_y": 0.0, "add": True}}],
        "numerics": _numerics(128, 20.0, 16.0, 1e-8),
        "accelerator": {"name": "anderson", "rank": 8, "step_size": 1.0},
        "misc": {"seed": seed},
    }


def d_plus_is_initial(d_component, ratio=0.2):
    """The d + is starting state from the d component: the s component is
    i * ratio * d (components in the order dx2-y2, s)."""
    return [d_component, 1j * ratio * d_component]


def parameter_sweep(count=48, seed=2024, ppxi=20.0):
    """A parameter sweep of `count` grains drawn from a seeded generator:
    each a disc of radius U[5, 20] xi0 or a square of side U[10, 40] xi0
    (coin flip), at `ppxi` points per xi0, external flux U[0, 4] flux quanta
    (no induced potential), T U[0.1, 0.9] Tc, s- or d_{x2-y2}-wave pairing
    (coin flip), 64 angles, cutoff 16, FP32-friendly criterion 1e-8."""
    import random
    rng = random.Random(seed)
    configs = []
    for _ in range(count):
        disc = rng.random() < 0.5
        size = rng.uniform(5.0, 20.0) if disc else rng.uniform(10.0, 40.0)
        flux = rng.uniform(0.0, 4.0)
        temperature = rng.uniform(0.1, 0.9)
        sym = "s" if rng.random() < 0.5 else "dx2-y2"
        geometry = ({"disc": {"radius": size, "center_x": 0.0, "center_y": 0.0,
                              "add": True}} if disc else
                    {"regular_polygon": {"num_edges": 4, "side_length": size,
                                         "rotation": 0.0, "center_x": 0.0,
                                         "center_y": 0.0, "add": True}})
        configs.append({
            "physics": {"temperature": temperature, "external_flux_quanta": flux,
                        "penetration_depth": -1.0, "gauge": "symmetric",
                        "charge_sign": -1,
                        "order_parameter": {sym: {"critical_temperature": 1.0,
                                                  "initial_noise_stddev": 0.01}}},
            "geometry": [geometry],
            "numerics": _numerics(64, ppxi, 16.0, 1e-8),
            "accelerator": {"name": "congacc"},
            "misc": {"seed": seed},
        })
    return configs


def postprocess_config(num_angles=720, num_energies=128, per_block=32):
    """The spectroscopy settings of the reference's examples
    (examples/*/postprocess_config.json): energies [0, 0.5] with broadening
    0.008, criterion 1e-4 in the L2 norm."""
    return {"spectroscopy": {"energy_min": 0.0, "energy_max": 0.5,
                             "energy_broadening": 0.008, "num_energies": num_energies},
            "numerics": {"convergence_criterion": 1e-4, "norm": "l2",
                         "num_energies_per_block": per_block,
                         "num_fermi_momenta": num_angles, "num_iterations_burnin": -1,
                         "num_iterations_max": 1000, "num_iterations_min": 0},
            "misc": {"data_format": "csv", "load_path": "", "save_path": ""}}


WORKLOADS = {"octagon_dwave": octagon_dwave, "vortex_disc": vortex_disc,
             "swave_disc": swave_disc, "dplus_is_square": dplus_is_square}
