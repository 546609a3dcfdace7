"""Simulation configs shared by the tests (simulation_config.json schema).

`geometry_internal` is an extension key: components in the reference's
internal [-0.5, 0.5] frame with an explicit grain width, the way the
reference's own C++ tests build their grains (e.g. tests/accelerators_test.cu
and tests/physics/current_conservation_test.cu construct DiscGeometry and
PolygonGeometry from the grain fraction directly).
"""
import copy
import math


def base(order_parameter=None, temperature=0.5, flux=0.0, lam=-1.0,
         angles=16, ppxi=3.0, cutoff=16.0, norm="linf", criterion=1e-5,
         accelerator="picard", rotation=0.0, gauge="symmetric",
         components=None, width=10.0, geometry=None, **acc):
    cfg = {
        "physics": {
            "temperature": temperature, "external_flux_quanta": flux,
            "penetration_depth": lam, "gauge": gauge, "charge_sign": -1,
            "crystal_axes_rotation": rotation,
            "order_parameter": order_parameter or {"s": {"critical_temperature": 1.0}},
        },
        "numerics": {"convergence_criterion": criterion, "energy_cutoff": cutoff,
                     "norm": norm, "num_fermi_momenta": angles,
                     "points_per_coherence_length": ppxi,
                     "vector_potential_error": 0.0},
        "accelerator": dict({"name": accelerator}, **acc),
    }
    if geometry is not None:
        cfg["geometry"] = geometry
    else:
        cfg["geometry_internal"] = {
            "grain_width": width,
            "components": components or [{"disc": {"radius": 0.5}}]}
    return cfg


def regular_polygon_unit(n):
    """Regular n-gon scaled to unit bounding-box extent, as
    tests/physics/current_conservation_test.cu:62-94 builds it."""
    step = 2.0 * math.pi / n
    off = 0.5 * step
    xs = [math.cos(step * i + off) for i in range(n)]
    ys = [math.sin(step * i + off) for i in range(n)]
    ext = max(max(xs) - min(xs), max(ys) - min(ys))
    radius = 1.0 / ext
    side = 2.0 * radius * math.sin(0.5 * step)
    return {"regular_polygon": {"num_edges": n, "side_length": side}}


# The accelerator known answer: tests/accelerators_test.cu:42-150.
def accelerator_case(name, **acc):
    return base(temperature=0.5, flux=1.0, lam=10.0, angles=16, ppxi=3.0,
                cutoff=16.0, norm="linf", criterion=1e-5, accelerator=name, **acc)


# A small grain with every feature on: two components, rotated crystal
# axes, field and induced vector potential, polygonal edges.
def mixed_small():
    return base(order_parameter={"dx2-y2": {"critical_temperature": 1.0},
                                 "s": {"critical_temperature": 0.3}},
                temperature=0.3, flux=0.7, lam=4.0, angles=24, ppxi=3.0,
                cutoff=12.0, rotation=0.0625,
                components=[regular_polygon_unit(8),
                            {"disc": {"radius": 0.15, "add": False,
                                      "center_x": 0.1, "center_y": -0.05}}])


def with_geometry(cfg, components, width=None):
    c = copy.deepcopy(cfg)
    c["geometry_internal"]["components"] = components
    if width is not None:
        c["geometry_internal"]["grain_width"] = width
    return c


def octagon_dwave(circumradius=25.0, ppxi=20.0, angles=128, temperature=0.1,
                  cutoff=32.0, lam=5.0):
    """The d-wave octagon: regular 8-gon of circumradius 25 xi0 (side
    2 R sin(pi/8)), d_{x2-y2} with crystal axes at 45 degrees, zero external
    field, self-consistent vector potential, Anderson with history 8."""
    side = 2.0 * circumradius * math.sin(math.pi / 8.0)
    return base(order_parameter={"dx2-y2": {"critical_temperature": 1.0}},
                temperature=temperature, flux=0.0, lam=lam, angles=angles,
                ppxi=ppxi, cutoff=cutoff, norm="l2", criterion=1e-8,
                accelerator="anderson", rank=8, step_size=1.0, rotation=0.125,
                geometry=[{"regular_polygon": {"num_edges": 8, "side_length": side,
                                               "rotation": 0.0}}])
