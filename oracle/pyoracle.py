"""ctypes view of the oracle (libsfo.so) and of the reference build (_ref/libsfref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by ``bench.py``'s cpu_baseline leg, never by the product package.

Both libraries export the same entry points under two prefixes: ``sfo_`` for the
restatement in ``sfo.cpp`` and ``sfr_`` for the reference components driven by
``ref_harness.cpp``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libsfo.so")
REF_SO = os.path.join(HERE, "_ref", "libsfref.so")
# the same sources under the reference's own Release flags
# (cmake/CompilerFlags.cmake:42,49; -march=x86-64-v3 in place of -march=native,
# since it is built on one host and timed on another): the CPU baseline, never
# a parity checker
REF_RELEASE_SO = os.path.join(HERE, "_ref", "libsfref_release.so")
REF_RELEASE_FLAGS = "-O3 -march=x86-64-v3 -ffast-math -fopenmp -DNDEBUG"


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("dt", C.c_double), ("ibm_iterations", C.c_int),
        ("tau", C.c_double), ("rho0", C.c_double), ("boundary_type", C.c_int),
        ("body_force_x", C.c_double), ("body_force_y", C.c_double),
        ("gravity_x", C.c_double), ("gravity_y", C.c_double),
        ("apply_gravity_to_fluid", C.c_int), ("apply_gravity_to_capsules", C.c_int),
        ("collision_model", C.c_int), ("use_mrt", C.c_int),
        ("mrt_se", C.c_double), ("mrt_s_eps", C.c_double), ("mrt_sq", C.c_double),
        ("periodic_y", C.c_int), ("max_lattice_force", C.c_double),
        ("rep_epsilon", C.c_double), ("rep_sigma", C.c_double), ("rep_r_cut", C.c_double),
        ("rep_power", C.c_int), ("rep_damping_normal", C.c_double),
        ("rep_friction_coeff", C.c_double),
        ("scalar_enabled", C.c_int), ("n_species", C.c_int),
        ("diffusivity", C.c_double * 4), ("inlet_concentration", C.c_double * 4),
        ("scalar_periodic_y", C.c_int),
        ("top_wall_velocity", C.c_double), ("bottom_wall_velocity", C.c_double),
        ("inlet_velocity", C.c_double), ("outlet_density", C.c_double),
        ("use_interpolated_bb", C.c_int),
        ("capsule_periodic_x", C.c_int),
    ]


class Membrane(C.Structure):
    _fields_ = [
        ("model", C.c_int), ("k_stretch", C.c_double), ("G_s", C.c_double),
        ("C_skalak", C.c_double), ("wlc_L_max_ratio", C.c_double),
        ("wlc_kBT_p", C.c_double), ("wlc_k_pow", C.c_double), ("k_bend", C.c_double),
        ("use_helfrich", C.c_int), ("kappa_0", C.c_double), ("k_area", C.c_double),
        ("k_perimeter", C.c_double), ("gamma_visc", C.c_double),
        ("eta_membrane", C.c_double), ("viscosity_ratio", C.c_double),
        ("density", C.c_double), ("shape", C.c_int), ("aspect_ratio", C.c_double),
        ("indent_depth", C.c_double), ("is_rigid", C.c_int),
    ]


_libs: dict = {}
_DP = C.POINTER(C.c_double)


def _lib(path: str, prefix: str):
    key = (path, prefix)
    if key in _libs:
        return _libs[key]
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run `make -C oracle`)")
    L = C.CDLL(path)
    g = lambda n: getattr(L, prefix + n)  # noqa: E731
    g("create").restype = C.c_void_p
    g("create").argtypes = [C.POINTER(Params)]
    g("destroy").argtypes = [C.c_void_p]
    g("add_capsule").argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int,
                                 C.POINTER(Membrane), C.c_int]
    g("add_capsule_random").restype = C.c_int
    g("add_capsule_random").argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_int, C.POINTER(Membrane), C.c_int, C.c_int,
                                        C.c_double, C.c_int]
    for n in ("initialize", "membrane_forces_only"):
        g(n).argtypes = [C.c_void_p]
    g("step").argtypes = [C.c_void_p, C.c_int]
    if prefix == "sfr_":
        g("set_threads").argtypes = [C.c_int]
        g("output_open").argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_int)]
        g("output_write").argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int]
        g("output_close").argtypes = [C.c_void_p]
    g("scalar_initialize").argtypes = [C.c_void_p, C.c_double]
    g("set_scalar_point").argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int]
    g("num_capsules").restype = C.c_int
    g("num_capsules").argtypes = [C.c_void_p]
    g("total_nodes").restype = C.c_int
    g("total_nodes").argtypes = [C.c_void_p]
    g("capsule_offsets").argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    g("get_nodes").argtypes = [C.c_void_p] + [_DP] * 6
    g("get_field").argtypes = [C.c_void_p] + [_DP] * 5
    g("get_f").argtypes = [C.c_void_p, _DP]
    g("get_scalar").argtypes = [C.c_void_p, _DP, C.c_int]
    g("last_capped").restype = C.c_int
    g("last_capped").argtypes = [C.c_void_p]
    g("capsule_geometry").restype = C.c_double
    g("capsule_geometry").argtypes = [C.c_void_p, C.c_int, C.c_int]
    g("phi").restype = C.c_double
    g("phi").argtypes = [C.c_double]
    g("set_nodes").argtypes = [C.c_void_p] + [_DP] * 4
    g("set_flags").argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    g("get_flags").argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    g("add_obstacle").argtypes = [C.c_void_p, C.c_int, _DP, C.c_int]
    g("obstacle_distance").restype = C.c_double
    g("obstacle_distance").argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, _DP]
    g("set_concentration").argtypes = [C.c_void_p, _DP, C.c_int]
    g("cap_forces").restype = C.c_int
    g("cap_forces").argtypes = [_DP, _DP, C.c_int, C.c_double]
    g("clamp_nodes_y").argtypes = [_DP, C.c_int, C.c_int]
    g("recirculate_x").restype = C.c_int
    g("recirculate_x").argtypes = [_DP, _DP, C.POINTER(C.c_int), C.c_int, C.c_int]
    if prefix == "sfo_":
        g("ibb_links").restype = C.c_int
        g("ibb_links").argtypes = [C.c_void_p, C.POINTER(C.c_int), _DP]
        for n in ("default_params",):
            g(n).argtypes = [C.POINTER(Params)]
        g("default_membrane").argtypes = [C.POINTER(Membrane)]
        g("set_f").argtypes = [C.c_void_p, _DP]
        g("set_macro").argtypes = [C.c_void_p, _DP, _DP, _DP]
        g("max_speed").restype = C.c_double
        g("max_speed").argtypes = [C.c_void_p]
        g("rest_lengths").argtypes = [C.c_void_p, _DP]
        g("phi_exact").restype = C.c_double
        g("phi_exact").argtypes = [C.c_double]
        g("pair_force").argtypes = [C.c_void_p, _DP, _DP, _DP, _DP, _DP]
        g("set_mdf_control").argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int]
        g("mdf_stats").argtypes = [C.c_void_p, _DP]
        g("mlp_num_params").restype = C.c_int
        g("mlp_num_params").argtypes = [C.c_int]
        g("mlp_predict").argtypes = [_DP, _DP, C.c_int, C.c_int, _DP, _DP]
        g("mlp_loss_grad").restype = C.c_double
        g("mlp_loss_grad").argtypes = [_DP, _DP, C.c_int, C.c_int, _DP, _DP, _DP]
        g("collide_cell").argtypes = [C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.c_double, _DP, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_double]
    _libs[key] = L
    return L


def oracle_lib():
    return _lib(ORACLE_SO, "sfo_")


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def _step_lib(which):
    return (oracle_lib(), "sfo_") if which == "oracle" else (_lib(REF_SO, "sfr_"), "sfr_")


def cap_forces(Fx, Fy, max_force, which="oracle"):
    """The step's force cap on copies of (Fx, Fy): (Fx', Fy', cells capped)."""
    L, pre = _step_lib(which)
    fx = np.ascontiguousarray(Fx, dtype=np.float64).copy()
    fy = np.ascontiguousarray(Fy, dtype=np.float64).copy()
    n = getattr(L, pre + "cap_forces")(_ptr(fx), _ptr(fy), fx.size, float(max_force))
    return fx, fy, n


def clamp_nodes_y(y, ny, which="oracle"):
    """The step's node clamp into [1.5, ny - 1.5] on a copy of y."""
    L, pre = _step_lib(which)
    yy = np.ascontiguousarray(y, dtype=np.float64).copy()
    getattr(L, pre + "clamp_nodes_y")(_ptr(yy), yy.size, int(ny))
    return yy


def recirculate_x(x, y, off, nx, which="oracle"):
    """The step's capsule recirculation in x on a copy of x: (x', capsules moved)."""
    L, pre = _step_lib(which)
    xx = np.ascontiguousarray(x, dtype=np.float64).copy()
    yy = np.ascontiguousarray(y, dtype=np.float64)
    oo = np.ascontiguousarray(off, dtype=np.int32)
    n = getattr(L, pre + "recirculate_x")(_ptr(xx), _ptr(yy), oo.ctypes.data_as(C.POINTER(C.c_int)),
                                          oo.size - 1, int(nx))
    return xx, n


def default_params(**kw) -> Params:
    p = Params()
    oracle_lib().sfo_default_params(C.byref(p))
    for k, v in kw.items():
        if k in ("diffusivity", "inlet_concentration"):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(p, k, v)
    return p


def default_membrane(**kw) -> Membrane:
    m = Membrane()
    oracle_lib().sfo_default_membrane(C.byref(m))
    for k, v in kw.items():
        setattr(m, k, v)
    return m


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_DP)


class Sim:
    """One simulation in either library (``which`` = "oracle" or "ref")."""

    def __init__(self, params: Params, which: str = "oracle"):
        if which == "oracle":
            self.L, self.pre = oracle_lib(), "sfo_"
        elif which == "ref":
            self.L, self.pre = _lib(REF_SO, "sfr_"), "sfr_"
        elif which == "ref-release":
            self.L, self.pre = _lib(REF_RELEASE_SO, "sfr_"), "sfr_"
        else:
            raise ValueError(which)
        self.params = params
        self.nx, self.ny = params.nx, params.ny
        self.h = self._f("create")(C.byref(params))

    def _f(self, name):
        return getattr(self.L, self.pre + name)

    def __del__(self):
        try:
            self._f("destroy")(self.h)
        except Exception:
            pass

    def add_capsule(self, cx, cy, r, nodes, mem: Membrane, typ=0):
        self._f("add_capsule")(self.h, cx, cy, r, nodes, C.byref(mem), typ)

    def add_capsule_random(self, count, x0, y0, x1, y1, rmin, rmax, nodes, mem, typ=0,
                           seed=12345, min_gap=1.0, max_attempts=0):
        return self._f("add_capsule_random")(self.h, count, x0, y0, x1, y1, rmin, rmax, nodes,
                                             C.byref(mem), typ, seed, min_gap, max_attempts)

    def add_obstacle(self, obstacle):
        """("circle", cx, cy, r) | ("rect", x0, y0, x1, y1) | ("polygon", [(x, y), ...])"""
        kind = {"circle": 0, "rect": 1, "polygon": 2}[obstacle[0]]
        vals = ([float(v) for xy in obstacle[1] for v in xy] if kind == 2
                else [float(v) for v in obstacle[1:]])
        a = np.asarray(vals, dtype=np.float64)
        self._f("add_obstacle")(self.h, kind, _ptr(a), a.size)

    def obstacle_distance(self, i, x, y):
        """(signedDistance, normalAt) of obstacle i at (x, y)."""
        nn = np.zeros(2)
        d = self._f("obstacle_distance")(self.h, i, x, y, _ptr(nn))
        return d, nn

    def ibb_links(self):
        """The interpolated bounce-back links (oracle only): int (L, 3) of (x, y, q), qf (L,)."""
        n = self.L.sfo_ibb_links(self.h, None, None)
        xyq = np.zeros(3 * max(n, 1), dtype=np.int32)
        qf = np.zeros(max(n, 1))
        self.L.sfo_ibb_links(self.h, xyq.ctypes.data_as(C.POINTER(C.c_int)), _ptr(qf))
        return xyq[:3 * n].reshape(n, 3), qf[:n]

    def initialize(self):
        self._f("initialize")(self.h)

    def scalar_initialize(self, c0):
        self._f("scalar_initialize")(self.h, c0)

    def set_scalar_point(self, x, y, c, sp=0):
        self._f("set_scalar_point")(self.h, x, y, c, sp)

    def step(self, n=1):
        self._f("step")(self.h, n)

    def set_threads(self, n):
        """OpenMP threads of the reference build's parallel regions (reference only)."""
        if self.pre != "sfr_":
            raise ValueError("the oracle restatement is single-threaded")
        self.L.sfr_set_threads(n)

    def output_open(self, directory, legacy=False, all_csv=False, extra_fluid=False, probes=()):
        """The reference's writers into `directory` (reference only)."""
        if self.pre != "sfr_":
            raise ValueError("the writers are the reference's own")
        ij = np.asarray([v for p in probes for v in p] or [0], dtype=np.int32)
        self._probe_ij = ij
        opts = (1 if legacy else 0) | (2 if all_csv else 0) | (4 if extra_fluid else 0)
        self.L.sfr_output_open(self.h, str(directory).encode(), opts, len(probes),
                               ij.ctypes.data_as(C.POINTER(C.c_int)))

    def output_write(self, step, time, vtk=False, csv=False, probes=False, stats=False):
        mask = (1 if vtk else 0) | (2 if csv else 0) | (4 if probes else 0) | (8 if stats else 0)
        self.L.sfr_output_write(self.h, step, time, mask)

    def output_close(self):
        self.L.sfr_output_close(self.h)

    def membrane_forces_only(self):
        self._f("membrane_forces_only")(self.h)

    @property
    def num_capsules(self):
        return self._f("num_capsules")(self.h)

    @property
    def total_nodes(self):
        return self._f("total_nodes")(self.h)

    def offsets(self):
        o = np.zeros(self.num_capsules + 1, dtype=np.int32)
        self._f("capsule_offsets")(self.h, o.ctypes.data_as(C.POINTER(C.c_int)))
        return o

    def nodes(self):
        K = self.total_nodes
        arrs = [np.zeros(K) for _ in range(6)]
        self._f("get_nodes")(self.h, *[_ptr(a) for a in arrs])
        return dict(zip(("x", "y", "vx", "vy", "fx", "fy"), arrs))

    def set_nodes(self, x=None, y=None, vx=None, vy=None):
        conv = lambda a: None if a is None else _ptr(np.ascontiguousarray(a, dtype=np.float64))  # noqa: E731
        keep = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None
                for a in (x, y, vx, vy)]
        self._f("set_nodes")(self.h, *[None if a is None else _ptr(a) for a in keep])
        del conv

    def field(self):
        N = self.nx * self.ny
        arrs = [np.zeros(N) for _ in range(5)]
        self._f("get_field")(self.h, *[_ptr(a) for a in arrs])
        return {k: a.reshape(self.ny, self.nx)
                for k, a in zip(("rho", "ux", "uy", "fx", "fy"), arrs)}

    def f(self):
        a = np.zeros(9 * self.nx * self.ny)
        self._f("get_f")(self.h, _ptr(a))
        return a.reshape(9, self.ny, self.nx)

    def set_f(self, f):
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        self._f("set_f")(self.h, _ptr(a))

    def set_macro(self, rho, ux, uy):
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (rho, ux, uy)]
        self._f("set_macro")(self.h, *[_ptr(a) for a in arrs])

    def flags(self):
        a = np.zeros(self.nx * self.ny, dtype=np.int32)
        self._f("get_flags")(self.h, a.ctypes.data_as(C.POINTER(C.c_int)))
        return a.reshape(self.ny, self.nx)

    def set_flags(self, flags):
        a = np.ascontiguousarray(flags, dtype=np.int32).ravel()
        assert a.size == self.nx * self.ny
        self._keep_flags = a
        self._f("set_flags")(self.h, a.ctypes.data_as(C.POINTER(C.c_int)))

    def set_mdf_control(self, tol, max_corrections, probe=False):
        self._f("set_mdf_control")(self.h, tol, max_corrections, 1 if probe else 0)

    def mdf_stats(self):
        out = np.zeros(4)
        self._f("mdf_stats")(self.h, _ptr(out))
        return out

    def scalar(self, sp=0):
        a = np.zeros(self.nx * self.ny)
        self._f("get_scalar")(self.h, _ptr(a), sp)
        return a.reshape(self.ny, self.nx)

    def set_concentration(self, C_, sp=0):
        a = np.ascontiguousarray(C_, dtype=np.float64).ravel()
        assert a.size == self.nx * self.ny
        self._f("set_concentration")(self.h, _ptr(a), sp)

    def last_capped(self):
        return self._f("last_capped")(self.h)

    def geometry(self, c, what):
        return self._f("capsule_geometry")(self.h, c, what)

    def max_speed(self):
        return self._f("max_speed")(self.h)

    def pair_force(self, pa, va, pb, vb):
        arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (pa, va, pb, vb)]
        out = np.zeros(2)
        self._f("pair_force")(self.h, *[_ptr(a) for a in arrs], _ptr(out))
        return out

    def rest_lengths(self):
        a = np.zeros(self.total_nodes)
        self._f("rest_lengths")(self.h, _ptr(a))
        return a


def mlp_predict(params, norm, x, width=64):
    """FP64 restatement of the warm-start network: y[n, 2] for x[n, 17]."""
    P = np.ascontiguousarray(params, dtype=np.float64)
    nr = np.ascontiguousarray(norm, dtype=np.float64)
    X = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 17)
    Y = np.zeros((X.shape[0], 2))
    oracle_lib().sfo_mlp_predict(_ptr(P), _ptr(nr), width, X.shape[0], _ptr(X), _ptr(Y))
    return Y


def mlp_loss_grad(params, norm, x, t, width=64):
    """(loss, dloss/dparams) of the warm-start network in FP64."""
    L = oracle_lib()
    P = np.ascontiguousarray(params, dtype=np.float64)
    nr = np.ascontiguousarray(norm, dtype=np.float64)
    X = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 17)
    T = np.ascontiguousarray(t, dtype=np.float64).reshape(-1, 2)
    G = np.zeros(L.sfo_mlp_num_params(width))
    loss = L.sfo_mlp_loss_grad(_ptr(P), _ptr(nr), width, X.shape[0], _ptr(X), _ptr(T), _ptr(G))
    return loss, G


def collide_cell(model, tau, f9, rho, ux, uy, Fx, Fy, se=1.4, seps=1.4, sq=1.4):
    f = np.ascontiguousarray(f9, dtype=np.float64).copy()
    oracle_lib().sfo_collide_cell(model, tau, se, seps, sq, _ptr(f), rho, ux, uy, Fx, Fy)
    return f


def phi(r):
    return oracle_lib().sfo_phi(r)


def phi_exact(r):
    return oracle_lib().sfo_phi_exact(r)
