"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, the smoke check and the CPU-baseline leg of bench.py may import
this module.  It builds oracle/_build/libconga_oracle.so on demand from
oracle/conga_oracle.cpp.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libconga_oracle.so")

MAXC = 4
MAXG = 16
DISC, REGULAR_POLYGON, POLYGON = 0, 1, 2
SYMMETRIES = {"s": 0, "dx2-y2": 1, "dxy": 2, "g": 3}
GAUGES = {"landau": 0, "symmetric": 1, "solenoid": 2}
NORMS = {"l1": 0, "l2": 1, "linf": 2}
ACCELERATORS = {"picard": 0, "anderson": 1, "congacc": 2,
                "barzilai-borwein": 3, "bb": 3, "polyak": 4, "broyden": 5}


class OcConfig(ctypes.Structure):
    _fields_ = [
        ("temperature", ctypes.c_double),
        ("external_flux_quanta", ctypes.c_double),
        ("penetration_depth", ctypes.c_double),
        ("crystal_axes_rotation", ctypes.c_double),
        ("gauge", ctypes.c_int),
        ("charge_sign", ctypes.c_int),
        ("num_components", ctypes.c_int),
        ("symmetry", ctypes.c_int * MAXC),
        ("critical_temperature", ctypes.c_double * MAXC),
        ("geometry_internal", ctypes.c_int),
        ("grain_width", ctypes.c_double),
        ("num_geometry", ctypes.c_int),
        ("geometry_type", ctypes.c_int * MAXG),
        ("geometry_add", ctypes.c_int * MAXG),
        ("geometry_param", (ctypes.c_double * 5) * MAXG),
        ("polygon_num_vertices", ctypes.c_int * MAXG),
        ("polygon_vertex_offset", ctypes.c_int * MAXG),
        ("polygon_vertices_x", ctypes.POINTER(ctypes.c_double)),
        ("polygon_vertices_y", ctypes.POINTER(ctypes.c_double)),
        ("convergence_criterion", ctypes.c_double),
        ("energy_cutoff", ctypes.c_double),
        ("norm", ctypes.c_int),
        ("num_fermi_momenta", ctypes.c_int),
        ("points_per_coherence_length", ctypes.c_double),
        ("vector_potential_error", ctypes.c_double),
        ("compute_current", ctypes.c_int),
        ("compute_free_energy", ctypes.c_int),
        ("accelerator", ctypes.c_int),
        ("step_size", ctypes.c_double),
        ("rank", ctypes.c_int),
        ("drag", ctypes.c_double),
        ("step_size_factor", ctypes.c_double),
        ("cos_similarity_min", ctypes.c_double),
        ("step_size_max", ctypes.c_double),
        ("step_size_min", ctypes.c_double),
        ("zero_stale", ctypes.c_int),
        ("num_threads", ctypes.c_int),
    ]


_lib = None


def build():
    src = os.path.join(HERE, "conga_oracle.cpp")
    if (not os.path.exists(LIB_PATH)
            or os.path.getmtime(LIB_PATH) < os.path.getmtime(src)):
        subprocess.run(["make", "-C", HERE], check=True,
                       stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        L.oc_create.restype = vp
        L.oc_create.argtypes = [ctypes.POINTER(OcConfig)]
        L.oc_destroy.argtypes = [vp]
        L.oc_last_error.restype = ctypes.c_char_p
        L.oc_get_int.restype = ctypes.c_longlong
        L.oc_get_int.argtypes = [vp, ctypes.c_char_p]
        L.oc_get_double.restype = ctypes.c_double
        L.oc_get_double.argtypes = [vp, ctypes.c_char_p]
        L.oc_get_array.restype = ctypes.c_longlong
        L.oc_get_array.argtypes = [vp, ctypes.c_char_p, dp, ctypes.c_longlong]
        L.oc_set_order_parameter.argtypes = [vp, ctypes.c_int, dp, dp]
        L.oc_set_vector_potential.argtypes = [vp, dp, dp]
        L.oc_set_array.argtypes = [vp, ctypes.c_char_p, dp, ctypes.c_longlong]
        L.oc_initial_bulk.argtypes = [vp, ctypes.c_double]
        L.oc_get_geometry.argtypes = [vp, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_byte), dp, dp, dp]
        L.oc_compute.argtypes = [vp]
        L.oc_burnin.argtypes = [vp]
        L.oc_run.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oc_ldos_setup.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_int]
        L.oc_ldos_compute.argtypes = [vp, ctypes.c_int]
        L.oc_ldos_run.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oc_free_energy_gradient.argtypes = [vp, dp, dp]
        L.oc_free_energy_gradient_a.argtypes = [vp, dp, dp, dp]
        L.oc_residual_gradient.argtypes = [vp, dp, dp, dp]
        L.oc_plain_map.argtypes = [vp, dp, dp]
        L.oc_sweep_angles.restype = ctypes.c_longlong
        L.oc_sweep_angles.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.oc_riccati_step.argtypes = [dp, dp, dp, ctypes.c_double, dp]
        L.oc_homogeneous.argtypes = [dp, dp, dp, dp]
        L.oc_greens.argtypes = [dp, dp, dp, dp]
        L.oc_num_ozaki.argtypes = [ctypes.c_double, ctypes.c_double]
        L.oc_ozaki.argtypes = [ctypes.c_int, dp, dp]
        L.oc_lattice.argtypes = [ctypes.c_double, ctypes.c_double,
                                 ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(ctypes.c_int), dp]
        for f in ("oc_bulk_gap",):
            getattr(L, f).restype = ctypes.c_double
            getattr(L, f).argtypes = [ctypes.c_double]
        L.oc_bulk_free_energy.restype = ctypes.c_double
        L.oc_bulk_free_energy.argtypes = [ctypes.c_double, ctypes.c_double]
        L.oc_yoshida.restype = ctypes.c_double
        L.oc_yoshida.argtypes = [ctypes.c_double] * 3
        L.oc_vector_potential.argtypes = [ctypes.c_int, ctypes.c_double,
                                          ctypes.c_int, ctypes.c_int, dp, dp]
        L.oc_accel_create.restype = ctypes.c_void_p
        L.oc_accel_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double]
        L.oc_accel_destroy.argtypes = [ctypes.c_void_p]
        L.oc_accel_step.argtypes = [ctypes.c_void_p, dp, dp]
        L.oc_congacc_inertia.restype = ctypes.c_double
        L.oc_congacc_inertia.argtypes = [ctypes.c_double] * 4
        L.oc_vector_potential_ex.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                             dp, dp]
        L.oc_poisson_kernel.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_config(cfg):
    """Map a simulation_config.json-style dict onto the oracle struct.

    Extra keys (not in the reference schema) accepted for the tests:
    cfg["geometry_internal"] = {"grain_width": w, "components": [...]}, with
    components {"disc": {...}} or {"regular_polygon": {...}} in the internal
    [-0.5, 0.5] frame; cfg["compute"] = {"current": bool, "free_energy": bool}.
    """
    c = OcConfig()
    ph = cfg["physics"]
    c.temperature = ph["temperature"]
    c.external_flux_quanta = ph.get("external_flux_quanta", 0.0)
    c.penetration_depth = ph.get("penetration_depth", -1.0)
    c.crystal_axes_rotation = 2.0 * math.pi * ph.get("crystal_axes_rotation", 0.0)
    c.gauge = GAUGES[ph.get("gauge", "symmetric")]
    c.charge_sign = ph.get("charge_sign", -1)
    comps = []
    # simulation.cu:201 adds the components in this order; the extension key
    # "component_order" overrides it (the reference's C++ tests add them in
    # any order, tests/free_energy_test.cu:107-117).
    order = cfg.get("component_order", ("dx2-y2", "dxy", "g", "s"))
    for name in order:
        if name in ph["order_parameter"]:
            comps.append((name, ph["order_parameter"][name]))
    c.num_components = len(comps)
    for i, (name, p) in enumerate(comps):
        c.symmetry[i] = SYMMETRIES[name]
        c.critical_temperature[i] = p.get("critical_temperature", 1.0)
    keep = []
    if "geometry_internal" in cfg:
        gi = cfg["geometry_internal"]
        c.geometry_internal = 1
        c.grain_width = gi["grain_width"]
        geo = gi["components"]
    else:
        geo = cfg["geometry"]
    c.num_geometry = len(geo)
    xs, ys = [], []
    for i, g in enumerate(geo):
        kind, p = next(iter(g.items()))
        c.geometry_add[i] = 1 if p.get("add", True) else 0
        if kind == "disc":
            c.geometry_type[i] = DISC
            c.geometry_param[i][0] = p.get("center_x", 0.0)
            c.geometry_param[i][1] = p.get("center_y", 0.0)
            c.geometry_param[i][2] = p.get("radius", 0.5)
        elif kind == "regular_polygon":
            c.geometry_type[i] = REGULAR_POLYGON
            c.geometry_param[i][0] = p.get("center_x", 0.0)
            c.geometry_param[i][1] = p.get("center_y", 0.0)
            c.geometry_param[i][2] = p["num_edges"]
            c.geometry_param[i][3] = p.get("rotation", 0.0)
            c.geometry_param[i][4] = p.get("side_length", -1.0)
        elif kind == "polygon":
            c.geometry_type[i] = POLYGON
            c.polygon_vertex_offset[i] = len(xs)
            c.polygon_num_vertices[i] = len(p["vertices_x"])
            xs += list(p["vertices_x"])
            ys += list(p["vertices_y"])
        else:
            raise ValueError(kind)
    vx = np.ascontiguousarray(xs if xs else [0.0], dtype=np.float64)
    vy = np.ascontiguousarray(ys if ys else [0.0], dtype=np.float64)
    keep += [vx, vy]
    c.polygon_vertices_x = _dp(vx)
    c.polygon_vertices_y = _dp(vy)
    nu = cfg["numerics"]
    c.convergence_criterion = nu.get("convergence_criterion", 1e-6)
    c.energy_cutoff = nu["energy_cutoff"]
    c.norm = NORMS[nu.get("norm", "l2")]
    c.num_fermi_momenta = nu["num_fermi_momenta"]
    c.points_per_coherence_length = nu["points_per_coherence_length"]
    c.vector_potential_error = nu.get("vector_potential_error", 0.0)
    comp = cfg.get("compute", {})
    c.compute_current = 1 if comp.get("current", True) else 0
    c.compute_free_energy = 1 if comp.get("free_energy", True) else 0
    ac = cfg.get("accelerator", {"name": "congacc"})
    c.accelerator = ACCELERATORS[ac["name"]]
    c.step_size = ac.get("step_size", -1.0)
    c.rank = ac.get("rank", -1)
    c.drag = ac.get("drag", -1.0)
    c.step_size_factor = ac.get("step_size_factor", -1.0)
    c.cos_similarity_min = ac.get("cos_similarity_min", -1.0)
    c.step_size_max = ac.get("step_size_max", -1.0)
    c.step_size_min = ac.get("step_size_min", -1.0)
    c.zero_stale = cfg.get("zero_stale", 0)
    c.num_threads = cfg.get("num_threads", 0)
    return c, keep


class Oracle:
    def __init__(self, cfg):
        L = lib()
        self._cfg, self._keep = make_config(cfg)
        self.h = L.oc_create(ctypes.byref(self._cfg))
        if not self.h:
            raise RuntimeError(L.oc_last_error().decode())
        self.N0 = self.get_int("N0")
        self.N1 = self.get_int("N1")
        self.Ne = self.get_int("num_energies")
        self.Na = self.get_int("num_angles")
        self.Nc = self.get_int("num_components")

    def __del__(self):
        if getattr(self, "h", None):
            lib().oc_destroy(self.h)
            self.h = None

    def get_int(self, k):
        return lib().oc_get_int(self.h, k.encode())

    def get(self, k):
        return lib().oc_get_double(self.h, k.encode())

    def array(self, name, size):
        out = np.empty(size, dtype=np.float64)
        n = lib().oc_get_array(self.h, name.encode(), _dp(out), size)
        if n < 0:
            raise RuntimeError(lib().oc_last_error().decode())
        return out[:n]

    def delta(self, new=False):
        n0 = self.N0 * self.N0
        a = self.array("delta_new" if new else "delta", self.Nc * 2 * n0)
        a = a.reshape(self.Nc, 2, self.N0, self.N0)
        return a[:, 0] + 1j * a[:, 1]

    def current(self):
        return self.array("current", 2 * self.N0 ** 2).reshape(2, self.N0, self.N0)

    def vector_potential(self, new=False):
        name = "vector_potential_new" if new else "vector_potential"
        return self.array(name, 2 * self.N0 ** 2).reshape(2, self.N0, self.N0)

    def domain(self):
        return self.array("domain", self.N0 ** 2).reshape(self.N0, self.N0)

    def boundary(self, in_out, direction):
        """Sparse boundary storage; in_out 0/1 = in/out, direction 0/1 =
        gamma/gamma-tilde (same convention as the product's C ABI)."""
        io = ("in", "out")[in_out]
        base = ("gamma", "gammatilde")[direction]
        cnt = self.get_int("num_storage_%s_%s" % (io, base)) * self.get_int("num_energies")
        a = self.array("boundary_%s_%s" % (io, base), 2 * cnt)
        return a[:cnt] + 1j * a[cnt:]

    def set_order_parameter(self, comp, values):
        v = np.ascontiguousarray(values, dtype=np.complex128)
        re = np.ascontiguousarray(v.real)
        im = np.ascontiguousarray(v.imag)
        lib().oc_set_order_parameter(self.h, comp, _dp(re), _dp(im))

    def set_vector_potential(self, ax, ay):
        ax = np.ascontiguousarray(ax, dtype=np.float64)
        ay = np.ascontiguousarray(ay, dtype=np.float64)
        lib().oc_set_vector_potential(self.h, _dp(ax), _dp(ay))

    def initial_bulk(self, phase=0.0):
        lib().oc_initial_bulk(self.h, phase)

    def geometry(self, k):
        n = self.N1 * self.N1
        lab = np.empty(n, dtype=np.int8)
        d = np.empty(n)
        nx = np.empty(n)
        ny = np.empty(n)
        lib().oc_get_geometry(self.h, k,
                              lab.ctypes.data_as(ctypes.POINTER(ctypes.c_byte)),
                              _dp(d), _dp(nx), _dp(ny))
        s = (self.N1, self.N1)
        return lab.reshape(s), d.reshape(s), nx.reshape(s), ny.reshape(s)

    def compute(self):
        lib().oc_compute(self.h)

    def burnin(self):
        lib().oc_burnin(self.h)

    def run(self, burnin, nmin, nmax):
        return bool(lib().oc_run(self.h, burnin, nmin, nmax))

    def sweep_angles(self, a0, a1):
        return lib().oc_sweep_angles(self.h, a0, a1)

    def plain_map(self):
        """G(u) of one plain iteration with the boundary in-values held
        fixed: ((Nc, N0, N0) complex order parameter, (2, N0, N0) vector
        potential in the setters' units)."""
        n = self.N0
        d = np.zeros(self.Nc * 2 * n * n)
        a = np.zeros(2 * n * n)
        if lib().oc_plain_map(self.h, _dp(d), _dp(a)) != 0:
            raise RuntimeError(lib().oc_last_error().decode())
        d = d.reshape(self.Nc, 2, n, n)
        return d[:, 0] + 1j * d[:, 1], a.reshape(2, n, n)

    def residual_gradient(self):
        """(R, grad, grad_a): the self-consistency residual
        R = 1/2 sum_inside |Delta' - Delta|^2 + 1/2 sum |A' - A|^2 of one
        plain iteration (boundary in-values fixed) and its gradient, grad
        as free_energy_gradient's, grad_a (2, N0, N0) per unit of A as
        set_vector_potential takes it."""
        n = self.N0
        g = np.zeros(self.Nc * 2 * n * n)
        ga = np.zeros(2 * n * n)
        r = np.zeros(1)
        if lib().oc_residual_gradient(self.h, _dp(g), _dp(ga), _dp(r)) != 0:
            raise RuntimeError(lib().oc_last_error().decode())
        g = g.reshape(self.Nc, 2, n, n)
        return float(r[0]), g[:, 0] + 1j * g[:, 1], ga.reshape(2, n, n)

    def free_energy_gradient(self, vector_potential=False):
        """(F, grad): the free energy one iteration reports from the current
        state and dF/dRe Delta_c + i dF/dIm Delta_c as an (Nc, N0, N0)
        complex array (vector potential and boundary in-values fixed).  With
        vector_potential, (F, grad, gradA): gradA the (2, N0, N0) dF/dA_x,
        dF/dA_y of the induced vector potential as set_vector_potential
        takes it."""
        n = self.N0
        g = np.zeros(self.Nc * 2 * n * n)
        f = np.zeros(1)
        if vector_potential:
            ga = np.zeros(2 * n * n)
            if lib().oc_free_energy_gradient_a(self.h, _dp(g), _dp(ga), _dp(f)) != 0:
                raise RuntimeError(lib().oc_last_error().decode())
        elif lib().oc_free_energy_gradient(self.h, _dp(g), _dp(f)) != 0:
            raise RuntimeError(lib().oc_last_error().decode())
        g = g.reshape(self.Nc, 2, n, n)
        if vector_potential:
            return float(f[0]), g[:, 0] + 1j * g[:, 1], ga.reshape(2, n, n)
        return float(f[0]), g[:, 0] + 1j * g[:, 1]

    # Keldysh spectroscopy (src/postprocess.cu)
    def ldos_setup(self, spectroscopy, per_block=-1):
        sp = spectroscopy
        self.num_ldos_energies = sp["num_energies"]
        if lib().oc_ldos_setup(self.h, sp["energy_min"], sp["energy_max"], sp["num_energies"],
                               sp["energy_broadening"], per_block) != 0:
            raise RuntimeError(lib().oc_last_error().decode())

    def ldos_compute(self, burnin=False):
        lib().oc_ldos_compute(self.h, 1 if burnin else 0)

    def ldos_run(self, burnin, nmin, nmax):
        return bool(lib().oc_ldos_run(self.h, burnin, nmin, nmax))

    def ldos(self):
        n = self.num_ldos_energies
        return self.array("ldos", n * self.N0 ** 2).reshape(n, self.N0, self.N0)


def postprocess_config(simulation, postprocess):
    """The oracle config of the postprocess context (src/postprocess.cu:
    118-330): physics and geometry of the simulation, angles, criterion and
    norm of the postprocess."""
    import copy
    cfg = copy.deepcopy(simulation)
    nu = postprocess["numerics"]
    cfg["numerics"]["num_fermi_momenta"] = nu["num_fermi_momenta"]
    cfg["numerics"]["convergence_criterion"] = nu["convergence_criterion"]
    cfg["numerics"]["norm"] = nu.get("norm", "l2")
    return cfg


def riccati_step(delta, energy, init, step):
    d = np.array([delta.real, delta.imag])
    e = np.array([energy.real, energy.imag])
    i = np.array([init.real, init.imag])
    out = np.empty(2)
    lib().oc_riccati_step(_dp(d), _dp(e), _dp(i), step, _dp(out))
    return complex(out[0], out[1])


def homogeneous(delta, energy):
    d = np.array([delta.real, delta.imag])
    e = np.array([energy.real, energy.imag])
    g = np.empty(2)
    gt = np.empty(2)
    lib().oc_homogeneous(_dp(d), _dp(e), _dp(g), _dp(gt))
    return complex(*g), complex(*gt)


def greens(gamma, gtilde):
    a = np.array([gamma.real, gamma.imag])
    b = np.array([gtilde.real, gtilde.imag])
    g = np.empty(2)
    f = np.empty(2)
    lib().oc_greens(_dp(a), _dp(b), _dp(g), _dp(f))
    return complex(*g), complex(*f)


def num_ozaki(temperature, cutoff):
    return lib().oc_num_ozaki(temperature, cutoff)


def ozaki(n):
    p = np.empty(n)
    r = np.empty(n)
    lib().oc_ozaki(n, _dp(p), _dp(r))
    return p, r


def lattice(width, ppxi):
    n0 = ctypes.c_int()
    n1 = ctypes.c_int()
    gf = ctypes.c_double()
    lib().oc_lattice(width, ppxi, ctypes.byref(n0), ctypes.byref(n1),
                     ctypes.byref(gf))
    return n0.value, n1.value, gf.value


def bulk_gap(t):
    return lib().oc_bulk_gap(t)


def bulk_free_energy(t, d):
    return lib().oc_bulk_free_energy(t, d)


def yoshida(t, d, cutoff):
    return lib().oc_yoshida(t, d, cutoff)


def vector_potential(n, h, num_grain, current, enforce=True, max_error=0.0,
                     analytical=True):
    j = np.ascontiguousarray(current, dtype=np.float64).reshape(-1)
    out = np.empty(2 * n * n)
    lib().oc_vector_potential_ex(n, h, num_grain, 1 if enforce else 0, float(max_error),
                                 1 if analytical else 0, _dp(j), _dp(out))
    return out.reshape(2, n, n)


def poisson_kernel(n, h, max_error=0.0, analytical=True):
    """(G[dy, dx], number of rank-one filters) of ComputeVectorPotential."""
    out = np.empty(n * n)
    k = lib().oc_poisson_kernel(n, h, float(max_error), 1 if analytical else 0, _dp(out))
    return out.reshape(n, n), k


class Accelerator:
    """The oracle's accelerator step on its own: vectors of num_fields
    fields of n0 * n0 values; step(x, g) returns the next iterate."""

    def __init__(self, name, n0, num_fields, rank=-1, step=-1.0):
        self.n = n0 * n0 * num_fields
        self.h = lib().oc_accel_create(ACCELERATORS[name], n0, num_fields, rank, step)

    def step(self, x, g):
        x = np.array(x, dtype=np.float64).reshape(-1)
        g = np.ascontiguousarray(g, dtype=np.float64).reshape(-1)
        assert x.size == self.n and g.size == self.n
        if lib().oc_accel_step(self.h, _dp(x), _dp(g)) != 0:
            raise RuntimeError(lib().oc_last_error().decode())
        return x

    def __del__(self):
        if getattr(self, "h", None):
            lib().oc_accel_destroy(self.h)
            self.h = None


def congacc_inertia(dn, mn, dot, min_cos=0.7):
    return lib().oc_congacc_inertia(dn, mn, dot, min_cos)
