"""superconga_amd -- MI355X (gfx950) self-consistency path of SuperConga.

Python mirror of the reference's simulation interface over the C ABI of
libsuperconga_amd.so (include/superconga_amd.h).  The heavy lifting is done
by hand-written HIP kernels; this module only moves configuration and host
arrays across the boundary.  It fails loudly when the library is missing:
there is no CPU fallback.
"""
import ctypes
import json
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libsuperconga_amd.so")

FP64 = 0
MIXED = 1
MIXED_REFINE = 2
_PRECISION = {"fp64": FP64, "mixed": MIXED, "mixed_refine": MIXED_REFINE,
              FP64: FP64, MIXED: MIXED, MIXED_REFINE: MIXED_REFINE}


class SupercongaError(RuntimeError):
    """Error reported by the C ABI (status code and message)."""

    def __init__(self, code, message):
        super().__init__("superconga_amd error %d: %s" % (code, message))
        self.code = code


class Dims(ctypes.Structure):
    _fields_ = [("n0", ctypes.c_int), ("n1", ctypes.c_int),
                ("num_energies", ctypes.c_int), ("num_angles", ctypes.c_int),
                ("num_components", ctypes.c_int),
                ("num_grain_sites", ctypes.c_int),
                ("num_in_gamma", ctypes.c_longlong),
                ("num_out_gamma", ctypes.c_longlong),
                ("num_in_gammatilde", ctypes.c_longlong),
                ("num_out_gammatilde", ctypes.c_longlong),
                ("traj_steps", ctypes.c_longlong),
                ("h", ctypes.c_double), ("grain_fraction", ctypes.c_double)]


class Results(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "residual", "residual_boundary", "residual_order_parameter",
        "residual_current", "residual_free_energy",
        "residual_vector_potential", "free_energy", "free_energy_per_area",
        "magnetic_free_energy", "magnetic_free_energy_per_area")] + [
        ("iterations", ctypes.c_int), ("converged", ctypes.c_int),
        ("last_sweep_ms", ctypes.c_double),
        ("last_iteration_ms", ctypes.c_double),
        ("mixed_iterations", ctypes.c_int)]


class NewtonReport(ctypes.Structure):
    _fields_ = [("newton_iterations", ctypes.c_int), ("krylov_iterations", ctypes.c_int),
                ("map_evaluations", ctypes.c_int), ("converged", ctypes.c_int),
                ("residual", ctypes.c_double),
                ("history", ctypes.POINTER(ctypes.c_double)),
                ("history_capacity", ctypes.c_int), ("history_length", ctypes.c_int)]


class LdosStatus(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "num_energies", "energies_per_block", "block_start", "block_count", "iterations",
        "angle_chunk", "partially_converged", "converged")] + [
        (n, ctypes.c_double) for n in (
            "residual", "residual_boundary", "residual_ldos", "last_sweep_ms",
            "last_iteration_ms")] + [("traj_steps", ctypes.c_longlong)]


_lib = None
_DP = ctypes.POINTER(ctypes.c_double)


def lib():
    """Loads libsuperconga_amd.so (built by superconga_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libsuperconga_amd.so is not built (%s); run "
                "python -m superconga_amd.build" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.sc_last_error.restype = ctypes.c_char_p
        L.sc_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
        L.sc_destroy.argtypes = [vp]
        L.sc_get_dims.argtypes = [vp, ctypes.POINTER(Dims)]
        L.sc_get_energies.argtypes = [vp, _DP, _DP]
        L.sc_set_order_parameter.argtypes = [vp, ctypes.c_int, _DP, _DP]
        L.sc_get_order_parameter.argtypes = [vp, ctypes.c_int, ctypes.c_int, _DP, _DP]
        L.sc_set_vector_potential.argtypes = [vp, _DP, _DP]
        L.sc_get_vector_potential.argtypes = [vp, ctypes.c_int, _DP, _DP]
        L.sc_get_current.argtypes = [vp, _DP, _DP]
        L.sc_get_free_energy_density.argtypes = [vp, _DP]
        L.sc_get_boundary.argtypes = [vp, ctypes.c_int, ctypes.c_int, _DP, _DP]
        L.sc_set_boundary.argtypes = [vp, ctypes.c_int, _DP, _DP]
        L.sc_compute.argtypes = [vp, ctypes.c_int]
        L.sc_burnin.argtypes = [vp, ctypes.c_int]
        L.sc_run.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                             ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.sc_get_results.argtypes = [vp, ctypes.POINTER(Results)]
        L.sc_riccati_sweep.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.sc_boundary_condition.argtypes = [vp, _DP]
        L.sc_integration_begin.argtypes = [vp, ctypes.c_int]
        L.sc_sweep_angles.argtypes = [vp] + [ctypes.c_int] * 4
        L.sc_integrate_angles.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.sc_integration_end.argtypes = [vp, ctypes.c_int]
        L.sc_num_fields.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.sc_accelerator_step.argtypes = [vp, _DP, _DP]
        L.sc_set_sweep_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int]
        ip = ctypes.POINTER(ctypes.c_int)
        L.sc_get_sweep_shape.argtypes = [vp, ctypes.c_int, ip, ip, ip]
        L.sc_autotune_sweep.argtypes = [vp, ctypes.c_int]
        L.sc_poisson_solve.argtypes = [ctypes.c_int, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_int, _DP, _DP]
        L.sc_poisson_solve_ex.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_double, _DP, _DP]
        L.sc_poisson_kernel.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _DP,
                                        ctypes.POINTER(ctypes.c_int)]
        L.sc_riccati_steps.argtypes = [ctypes.c_int, _DP, _DP, ctypes.c_int]
        L.sc_num_ozaki.argtypes = [ctypes.c_double, ctypes.c_double]
        L.sc_ozaki.argtypes = [ctypes.c_int, _DP, _DP]
        L.sc_bulk_gap.restype = ctypes.c_double
        L.sc_bulk_gap.argtypes = [ctypes.c_double]
        L.sc_get_labels.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_byte)]
        L.sc_get_domain.argtypes = [vp, ctypes.POINTER(ctypes.c_byte)]
        L.sc_batch_create.argtypes = [ctypes.POINTER(ctypes.c_char_p), ctypes.c_int,
                                      ctypes.POINTER(vp)]
        L.sc_batch_destroy.argtypes = [vp]
        L.sc_batch_size.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.sc_batch_grain.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
        L.sc_batch_compute.argtypes = [vp, ctypes.c_int]
        L.sc_batch_burnin.argtypes = [vp, ctypes.c_int]
        L.sc_batch_run.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.sc_batch_set_sweep_shape.argtypes = [vp] + [ctypes.c_int] * 4
        L.sc_batch_get_sweep_shape.argtypes = [vp, ctypes.c_int] + [
            ctypes.POINTER(ctypes.c_int)] * 3
        L.sc_batch_autotune.argtypes = [vp, ctypes.c_int]
        L.sc_batch_timing.argtypes = [vp, _DP, _DP, ctypes.POINTER(ctypes.c_longlong)]
        L.sc_describe.argtypes = [ctypes.c_char_p, ctypes.c_int,
                                  ctypes.POINTER(Dims), ctypes.POINTER(ctypes.c_byte)]
        L.sc_newton_krylov.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                       ctypes.POINTER(NewtonReport)]
        L.sc_newton_krylov_ex.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_int, ctypes.POINTER(NewtonReport)]
        L.sc_set_graphs.argtypes = [vp, ctypes.c_int]
        L.sc_set_screening.argtypes = [vp, ctypes.c_double]
        L.sc_batch_newton_krylov.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(NewtonReport)]
        L.sc_batch_set_graphs.argtypes = [vp, ctypes.c_int]
        L.sc_free_energy_gradient.argtypes = [vp, _DP, _DP]
        L.sc_free_energy_gradient_ex.argtypes = [vp, _DP, _DP, _DP]
        L.sc_residual_gradient.argtypes = [vp, _DP, _DP, _DP]
        L.sc_jacobian_vector.argtypes = [vp, _DP, _DP]
        L.sc_set_adjoint_checkpoints.argtypes = [vp, ctypes.c_int]
        L.sc_set_newton_options.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                            ctypes.c_int]
        L.sc_ldos_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(vp)]
        L.sc_ldos_compute.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.sc_ldos_run.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_int)]
        L.sc_ldos_get_status.argtypes = [vp, ctypes.POINTER(LdosStatus)]
        L.sc_ldos_get.argtypes = [vp, _DP]
        L.sc_ldos_set_angle_chunk.argtypes = [vp, ctypes.c_int]
        _lib = L
    return _lib


def _check(code):
    if code != 0:
        raise SupercongaError(code, lib().sc_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(_DP)


# Jacobian-vector products of Newton-Krylov: finite differences on the FP64
# or the FP32-stepping map, or exact by the tangent-linear sweep
_JVP = {"fp64": FP64, "mixed": MIXED, "mixed_refine": MIXED_REFINE, "tangent": 3}


def _jvp(p):
    try:
        return _JVP[p]
    except KeyError:
        raise ValueError("jvp_precision must be 'fp64', 'mixed' or 'tangent'") from None


def _precision(p):
    try:
        return _PRECISION[p]
    except KeyError:
        raise ValueError("precision must be 'fp64', 'mixed' or 'mixed_refine'") from None


class Simulation:
    """One grain: Context<T> with its modules (Context.h:43-183)."""

    def __init__(self, config=None, _borrow=None):
        if _borrow is not None:
            self._h, self._owner = _borrow
            self._borrowed = True
        else:
            text = config if isinstance(config, str) else json.dumps(config)
            h = ctypes.c_void_p()
            _check(lib().sc_create(text.encode(), ctypes.byref(h)))
            self._h = h
            self._borrowed = False
        d = Dims()
        _check(lib().sc_get_dims(self._h, ctypes.byref(d)))
        self.dims = d
        self.N0, self.N1 = d.n0, d.n1
        self.Ne, self.Na, self.Nc = d.num_energies, d.num_angles, d.num_components

    def close(self):
        if getattr(self, "_h", None):
            if not getattr(self, "_borrowed", False):
                lib().sc_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ----- state I/O -----
    def energies(self):
        e = np.empty(self.Ne)
        r = np.empty(self.Ne)
        _check(lib().sc_get_energies(self._h, _dp(e), _dp(r)))
        return e, r

    def set_order_parameter(self, comp, values):
        v = np.asarray(values, dtype=np.complex128).reshape(self.N0, self.N0)
        re = np.ascontiguousarray(v.real)
        im = np.ascontiguousarray(v.imag)
        _check(lib().sc_set_order_parameter(self._h, comp, _dp(re), _dp(im)))

    def order_parameter(self, which=0):
        out = np.empty((self.Nc, self.N0, self.N0), dtype=np.complex128)
        re = np.empty((self.N0, self.N0))
        im = np.empty((self.N0, self.N0))
        for c in range(self.Nc):
            _check(lib().sc_get_order_parameter(self._h, which, c, _dp(re), _dp(im)))
            out[c] = re + 1j * im
        return out

    def set_vector_potential(self, ax, ay):
        ax = np.ascontiguousarray(ax, dtype=np.float64)
        ay = np.ascontiguousarray(ay, dtype=np.float64)
        _check(lib().sc_set_vector_potential(self._h, _dp(ax), _dp(ay)))

    def vector_potential(self, which=0):
        ax = np.empty((self.N0, self.N0))
        ay = np.empty((self.N0, self.N0))
        _check(lib().sc_get_vector_potential(self._h, which, _dp(ax), _dp(ay)))
        return np.stack([ax, ay])

    def current(self):
        jx = np.empty((self.N0, self.N0))
        jy = np.empty((self.N0, self.N0))
        _check(lib().sc_get_current(self._h, _dp(jx), _dp(jy)))
        return np.stack([jx, jy])

    def free_energy_density(self):
        out = np.empty((self.N0, self.N0))
        _check(lib().sc_get_free_energy_density(self._h, _dp(out)))
        return out

    def boundary(self, in_out, direction):
        d = self.dims
        n = {(0, 0): d.num_in_gamma, (1, 0): d.num_out_gamma,
             (0, 1): d.num_in_gammatilde, (1, 1): d.num_out_gammatilde}[
            (in_out, direction)] * self.Ne
        re = np.empty(n)
        im = np.empty(n)
        _check(lib().sc_get_boundary(self._h, in_out, direction, _dp(re), _dp(im)))
        return re + 1j * im

    def set_boundary(self, direction, values):
        v = np.asarray(values, dtype=np.complex128)
        re = np.ascontiguousarray(v.real)
        im = np.ascontiguousarray(v.imag)
        _check(lib().sc_set_boundary(self._h, direction, _dp(re), _dp(im)))

    def domain(self):
        """Reference-frame domain: 1 inside the grain, 0 outside."""
        out = np.empty(self.N0 * self.N0, dtype=np.int8)
        _check(lib().sc_get_domain(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_byte))))
        return out.reshape(self.N0, self.N0)

    def labels(self, angle):
        out = np.empty(self.N1 * self.N1, dtype=np.int8)
        _check(lib().sc_get_labels(self._h, angle,
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_byte))))
        return out.reshape(self.N1, self.N1)

    # ----- iteration -----
    def compute(self, precision="fp64"):
        _check(lib().sc_compute(self._h, _precision(precision)))

    def burnin(self, precision="fp64"):
        _check(lib().sc_burnin(self._h, _precision(precision)))

    def run(self, burnin=-1, nmin=-1, nmax=-1, precision="fp64"):
        c = ctypes.c_int()
        _check(lib().sc_run(self._h, burnin, nmin, nmax, _precision(precision),
                            ctypes.byref(c)))
        return bool(c.value)

    def results(self):
        r = Results()
        _check(lib().sc_get_results(self._h, ctypes.byref(r)))
        return {n: getattr(r, n) for n, _ in Results._fields_}

    def newton_krylov(self, max_newton=20, krylov_dim=20, tol=1e-9, jvp_precision="fp64"):
        """Jacobian-free Newton-Krylov on the fixed point of one plain FP64
        iteration; returns the report as a dict with the weighted residual
        history.  jvp_precision "mixed" takes the Jacobian-vector products
        on the map with FP32 stepping (the residuals stay FP64), "tangent"
        exactly by the tangent-linear sweep.  A
        following compute() reports the usual residuals."""
        hist = (ctypes.c_double * (max_newton + 1))()
        r = NewtonReport(history=ctypes.cast(hist, ctypes.POINTER(ctypes.c_double)),
                         history_capacity=max_newton + 1)
        _check(lib().sc_newton_krylov_ex(self._h, max_newton, krylov_dim, ctypes.c_double(tol),
                                         _jvp(jvp_precision), ctypes.byref(r)))
        return {"newton_iterations": r.newton_iterations,
                "krylov_iterations": r.krylov_iterations,
                "map_evaluations": r.map_evaluations, "converged": bool(r.converged),
                "residual": r.residual,
                "history": [hist[i] for i in range(r.history_length)]}

    def free_energy_gradient(self, vector_potential=False):
        """(F, grad): the free energy one iteration reports from the current
        state and its gradient dF/dRe Delta_c + i dF/dIm Delta_c, an
        (Nc, N0, N0) complex array, by the discrete adjoint of the iteration
        (vector potential and boundary in-values held fixed; the state is
        left as it was).  With vector_potential, (F, grad, grad_a): grad_a
        the (2, N0, N0) dF/dA_x, dF/dA_y of the induced vector potential per
        unit of A as set_vector_potential takes it."""
        n = self.N0
        g = np.empty(self.Nc * 2 * n * n)
        f = np.zeros(1)
        if vector_potential:
            ga = np.empty(2 * n * n)
            _check(lib().sc_free_energy_gradient_ex(self._h, _dp(g), _dp(ga), _dp(f)))
        else:
            _check(lib().sc_free_energy_gradient(self._h, _dp(g), _dp(f)))
        g = g.reshape(self.Nc, 2, n, n)
        if vector_potential:
            return float(f[0]), g[:, 0] + 1j * g[:, 1], ga.reshape(2, n, n)
        return float(f[0]), g[:, 0] + 1j * g[:, 1]

    def residual_gradient(self):
        """(R, grad, grad_a): the self-consistency residual
        R = 1/2 sum_inside |Delta' - Delta|^2 + 1/2 sum |A' - A|^2 of one
        plain FP64 iteration u -> G(u) (boundary in-values held fixed) and
        its gradient (J - I)^T (G(u) - u): grad an (Nc, N0, N0) complex
        array (dR/dRe + i dR/dIm), grad_a (2, N0, N0) in A as
        set_vector_potential takes it.  The state is left as it was."""
        n = self.N0
        g = np.empty(self.Nc * 2 * n * n)
        ga = np.empty(2 * n * n)
        r = np.zeros(1)
        _check(lib().sc_residual_gradient(self._h, _dp(g), _dp(ga), _dp(r)))
        g = g.reshape(self.Nc, 2, n, n)
        return float(r[0]), g[:, 0] + 1j * g[:, 1], ga.reshape(2, n, n)

    def set_newton_options(self, restarts=4, eta_max=0.3, fp64_fallback=True, fix_phase=False):
        """Newton-Krylov's GMRES cycles per Newton step, the cap of its
        Eisenstat-Walker forcing term, whether FP32 products hand over to
        FP64 ones when a line search finds no descent, and whether each
        Newton step drops its component along the order parameter's global
        phase rotation (a symmetry, so a singular direction of J - I)."""
        _check(lib().sc_set_newton_options(self._h, int(restarts), ctypes.c_double(eta_max),
                                           int(bool(fp64_fallback)), int(bool(fix_phase))))

    def set_adjoint_checkpoints(self, k):
        """The reverse sweeps' storage: k > 0 a checkpoint every k sites
        (each segment integrated again), -1 (the default) picks
        k = ceil(sqrt(longest trajectory)), 0 keeps every site.  The
        gradients are the same."""
        _check(lib().sc_set_adjoint_checkpoints(self._h, int(k)))

    def jacobian_vector(self, v_delta, v_a=None):
        """J v of one plain FP64 iteration u -> G(u) at the current state
        (boundary in-values held fixed), exactly by the tangent-linear
        sweep: v_delta an (Nc, N0, N0) complex array, v_a (2, N0, N0) in the
        units set_vector_potential takes (or None); returns
        (jv_delta, jv_a) alike."""
        n = self.N0
        v = np.zeros(self.Nc * 2 * n * n + 2 * n * n)
        vd = np.asarray(v_delta, dtype=np.complex128).reshape(self.Nc, n, n)
        for c in range(self.Nc):
            v[(2 * c) * n * n:(2 * c + 1) * n * n] = vd[c].real.ravel()
            v[(2 * c + 1) * n * n:(2 * c + 2) * n * n] = vd[c].imag.ravel()
        if v_a is not None:
            v[self.Nc * 2 * n * n:] = np.asarray(v_a, dtype=np.float64).ravel()
        jv = np.empty_like(v)
        _check(lib().sc_jacobian_vector(self._h, _dp(v), _dp(jv)))
        jd = jv[:self.Nc * 2 * n * n].reshape(self.Nc, 2, n, n)
        return jd[:, 0] + 1j * jd[:, 1], jv[self.Nc * 2 * n * n:].reshape(2, n, n)

    def riccati_sweep(self, precision="fp64", burnin=False):
        _check(lib().sc_riccati_sweep(self._h, _precision(precision), int(burnin)))

    # ----- the iteration angle by angle (IntegrationIteratorOzaki::compute) -----
    def integration_begin(self, burnin=False):
        _check(lib().sc_integration_begin(self._h, int(burnin)))

    def sweep_angles(self, begin, end, precision="fp64", burnin=False):
        _check(lib().sc_sweep_angles(self._h, _precision(precision), int(burnin), begin, end))

    def integrate_angles(self, begin, end):
        _check(lib().sc_integrate_angles(self._h, begin, end))

    def integration_end(self, burnin=False):
        _check(lib().sc_integration_end(self._h, int(burnin)))

    def num_fields(self):
        n = ctypes.c_int()
        _check(lib().sc_num_fields(self._h, ctypes.byref(n)))
        return n.value

    def accelerator_step(self, x, g):
        """Accelerator<T>::compute on caller-supplied fields: x (the
        iterate) and g (the map's result), each (num_fields, N0, N0);
        returns the accelerated iterate."""
        shape = (self.num_fields(), self.N0, self.N0)
        x = np.array(x, dtype=np.float64, order="C").reshape(shape)
        g = np.ascontiguousarray(g, dtype=np.float64).reshape(shape)
        _check(lib().sc_accelerator_step(self._h, _dp(x), _dp(g)))
        return x

    def boundary_condition(self):
        r = ctypes.c_double()
        _check(lib().sc_boundary_condition(self._h, ctypes.byref(r)))
        return r.value

    def set_sweep_shape(self, energies_per_lane, waves_per_simd, slots_per_cu,
                        precision="fp64"):
        _check(lib().sc_set_sweep_shape(self._h, _precision(precision), energies_per_lane,
                                        waves_per_simd, slots_per_cu))

    def set_graphs(self, enable=True):
        """Replay the stages after the sweep from a captured hipGraph (the
        default) or launch them eagerly."""
        _check(lib().sc_set_graphs(self._h, 1 if enable else 0))

    def set_screening(self, superfluid_density):
        """The screened vector-potential update on (superfluid density > 0)
        or off (0); the accelerator restarts."""
        _check(lib().sc_set_screening(self._h, ctypes.c_double(superfluid_density)))

    def sweep_shape(self, precision="fp64"):
        e, w, s = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().sc_get_sweep_shape(self._h, _precision(precision), ctypes.byref(e),
                                        ctypes.byref(w), ctypes.byref(s)))
        return e.value, w.value, s.value

    def autotune(self, precision="fp64"):
        """Picks the fastest sweep variant for this grain and precision."""
        _check(lib().sc_autotune_sweep(self._h, _precision(precision)))
        return self.sweep_shape(precision)


class Spectroscopy(Simulation):
    """The postprocess context of src/postprocess.cu: the LDOS at real
    energies E + i delta of a given order parameter and vector potential
    (set them with set_order_parameter / set_vector_potential)."""

    def __init__(self, simulation_config, postprocess_config):
        sim = simulation_config if isinstance(simulation_config, str) else json.dumps(
            simulation_config)
        post = postprocess_config if isinstance(postprocess_config, str) else json.dumps(
            postprocess_config)
        h = ctypes.c_void_p()
        _check(lib().sc_ldos_create(sim.encode(), post.encode(), ctypes.byref(h)))
        self.postprocess = json.loads(post)
        super().__init__(_borrow=(h, None))
        self._borrowed = False  # owned: destroyed on close()

    def compute(self, precision="fp64", burnin=False):
        _check(lib().sc_ldos_compute(self._h, _precision(precision), int(bool(burnin))))

    def burnin(self, precision="fp64"):
        self.compute(precision, burnin=True)

    def run(self, burnin=None, nmin=None, nmax=None, precision="fp64"):
        """runPostprocess; limits default to the postprocess config's."""
        nu = self.postprocess.get("numerics", {})
        burnin = nu.get("num_iterations_burnin", -1) if burnin is None else burnin
        nmin = nu.get("num_iterations_min", 0) if nmin is None else nmin
        nmax = nu.get("num_iterations_max", -1) if nmax is None else nmax
        c = ctypes.c_int()
        _check(lib().sc_ldos_run(self._h, burnin, nmin, nmax, _precision(precision),
                                 ctypes.byref(c)))
        return bool(c.value)

    def status(self):
        st = LdosStatus()
        _check(lib().sc_ldos_get_status(self._h, ctypes.byref(st)))
        return {n: getattr(st, n) for n, _ in LdosStatus._fields_}

    def ldos(self):
        n = self.status()["num_energies"]
        out = np.empty((n, self.N0, self.N0))
        _check(lib().sc_ldos_get(self._h, _dp(out)))
        return out

    def set_angle_chunk(self, angles):
        _check(lib().sc_ldos_set_angle_chunk(self._h, angles))

    def energies_ldos(self):
        """The real energies of the whole grid (IntegrationIteratorKeldysh
        initEnergies)."""
        sp = self.postprocess["spectroscopy"]
        n = sp["num_energies"]
        step = (sp["energy_max"] - sp["energy_min"]) / (n - 1) if n > 1 else 0.0
        return sp["energy_min"] + step * np.arange(n)


class Batch:
    """Many grains iterated together; one sweep launch per iteration covers
    every grain still iterating.  grain(i) gives the per-grain accessors of
    Simulation (borrowed: they live as long as the batch)."""

    def __init__(self, configs):
        texts = [c if isinstance(c, str) else json.dumps(c) for c in configs]
        arr = (ctypes.c_char_p * len(texts))(*[t.encode() for t in texts])
        h = ctypes.c_void_p()
        _check(lib().sc_batch_create(arr, len(texts), ctypes.byref(h)))
        self._h = h
        self._grains = []
        for i in range(len(texts)):
            g = ctypes.c_void_p()
            _check(lib().sc_batch_grain(self._h, i, ctypes.byref(g)))
            self._grains.append(Simulation(_borrow=(g, self)))

    def __len__(self):
        return len(self._grains)

    def grain(self, i):
        return self._grains[i]

    def close(self):
        if getattr(self, "_h", None):
            for g in self._grains:
                g._h = None
            lib().sc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def compute(self, precision="fp64"):
        _check(lib().sc_batch_compute(self._h, _precision(precision)))

    def burnin(self, precision="fp64"):
        _check(lib().sc_batch_burnin(self._h, _precision(precision)))

    def run(self, burnin=-1, nmin=-1, nmax=-1, precision="fp64"):
        conv = (ctypes.c_int * len(self._grains))()
        _check(lib().sc_batch_run(self._h, burnin, nmin, nmax, _precision(precision), conv))
        return [bool(c) for c in conv]

    def set_sweep_shape(self, energies_per_lane, waves_per_simd, slots_per_cu,
                        precision="fp64"):
        _check(lib().sc_batch_set_sweep_shape(self._h, _precision(precision),
                                              energies_per_lane, waves_per_simd,
                                              slots_per_cu))

    def set_graphs(self, enable=True):
        """Captured hipGraphs (the default) or eager launches for the stages
        after the sweep of every grain."""
        _check(lib().sc_batch_set_graphs(self._h, 1 if enable else 0))

    def newton_krylov(self, max_newton=20, krylov_dim=20, tol=1e-9, jvp_precision="fp64",
                      grains=None):
        """Newton-Krylov on many grains at once (Simulation.newton_krylov per
        grain, the maps of all solving grains in one sweep launch per
        precision); `grains` (indices, default all) selects the grains to
        solve.  One report dict per grain (zeros for grains not solved)."""
        n = len(self._grains)
        mask = None
        if grains is not None:
            sel = set(int(i) for i in grains)
            mask = (ctypes.c_int * n)(*[1 if i in sel else 0 for i in range(n)])
        hists = [(ctypes.c_double * (max_newton + 1))() for _ in range(n)]
        reps = (NewtonReport * n)()
        for i in range(n):
            reps[i].history = ctypes.cast(hists[i], ctypes.POINTER(ctypes.c_double))
            reps[i].history_capacity = max_newton + 1
        _check(lib().sc_batch_newton_krylov(self._h, max_newton, krylov_dim,
                                            ctypes.c_double(tol), _jvp(jvp_precision),
                                            mask, reps))
        return [{"newton_iterations": r.newton_iterations,
                 "krylov_iterations": r.krylov_iterations,
                 "map_evaluations": r.map_evaluations, "converged": bool(r.converged),
                 "residual": r.residual,
                 "history": [hists[i][q] for q in range(r.history_length)]}
                for i, r in enumerate(reps)]

    def sweep_shape(self, precision="fp64"):
        e, w, s = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().sc_batch_get_sweep_shape(self._h, _precision(precision),
                                              ctypes.byref(e), ctypes.byref(w),
                                              ctypes.byref(s)))
        return e.value, w.value, s.value

    def autotune(self, precision="fp64"):
        _check(lib().sc_batch_autotune(self._h, _precision(precision)))
        return self.sweep_shape(precision)

    def timing(self):
        """(sweep ms, iteration ms, trajectory steps) of the last iteration."""
        a, b = ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_longlong()
        _check(lib().sc_batch_timing(self._h, ctypes.byref(a), ctypes.byref(b),
                                     ctypes.byref(n)))
        return a.value, b.value, n.value


def describe(config, angle=None):
    """Host-only plan of a config: its sizes and, for one angle, its labels.

    Needs no device; returns (dims, labels or None)."""
    text = config if isinstance(config, str) else json.dumps(config)
    d = Dims()
    if angle is None:
        _check(lib().sc_describe(text.encode(), 0, ctypes.byref(d), None))
        return d, None
    # sizes first, then the labels buffer of the right size
    _check(lib().sc_describe(text.encode(), 0, ctypes.byref(d), None))
    out = np.empty(d.n1 * d.n1, dtype=np.int8)
    _check(lib().sc_describe(text.encode(), angle, ctypes.byref(d),
                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_byte))))
    return d, out.reshape(d.n1, d.n1)


def poisson_solve(current, h, num_grain, enforce_mean=True, max_error=0.0):
    """ComputeVectorPotential::solve: the exact Green function
    (max_error = 0) or its eigen-truncated expansion (max_error > 0, the
    config's vector_potential_error)."""
    j = np.ascontiguousarray(current, dtype=np.float64)
    n = j.shape[-1]
    out = np.empty((2, n, n))
    _check(lib().sc_poisson_solve_ex(n, h, num_grain, int(enforce_mean), float(max_error),
                                     _dp(j), _dp(out)))
    return out


def poisson_kernel(n, h, max_error=0.0):
    """(G, filters): the kernel G[dy, dx] the vector-potential solve
    convolves with, and the number of rank-one filters of its truncation
    (host only; no device needed)."""
    out = np.empty((n, n))
    k = ctypes.c_int()
    _check(lib().sc_poisson_kernel(n, h, float(max_error), _dp(out), ctypes.byref(k)))
    return out, k.value


def riccati_steps(delta, energy, initial, step, precision="fp64"):
    """Midpoint Riccati steps on the device for arrays of inputs."""
    d = np.atleast_1d(np.asarray(delta, dtype=np.complex128))
    e = np.broadcast_to(np.asarray(energy, dtype=np.complex128), d.shape)
    y = np.broadcast_to(np.asarray(initial, dtype=np.complex128), d.shape)
    s = np.broadcast_to(np.asarray(step, dtype=np.float64), d.shape)
    rec = np.ascontiguousarray(np.stack([d.real, d.imag, e.real, e.imag, y.real,
                                         y.imag, s], axis=-1).reshape(-1, 7))
    out = np.empty((rec.shape[0], 2))
    _check(lib().sc_riccati_steps(rec.shape[0], _dp(rec), _dp(out), _precision(precision)))
    return (out[:, 0] + 1j * out[:, 1]).reshape(d.shape)


def num_ozaki(temperature, cutoff):
    return lib().sc_num_ozaki(temperature, cutoff)


def ozaki(n):
    p = np.empty(n)
    r = np.empty(n)
    _check(lib().sc_ozaki(n, _dp(p), _dp(r)))
    return p, r


def bulk_gap(temperature):
    return lib().sc_bulk_gap(temperature)
