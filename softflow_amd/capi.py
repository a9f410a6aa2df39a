"""ctypes binding of the C-ABI in include/softflow.h (libsoftflow_hip.so).

This is the binding a Python caller of the reference engine would use to reach
the MI355X kernels directly, without the pybind engine.  Loading fails loudly
if the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libsoftflow_hip.so")

SF_OK, SF_E_ARG, SF_E_HIP, SF_E_STATE, SF_E_CAPACITY = 0, -1, -2, -3, -4
SF_FP64, SF_FP32 = 0, 1
SF_BGK, SF_MRT, SF_REGULARIZED = 0, 1, 2
SF_INLET_OUTLET, SF_PERIODIC, SF_CLOSED = 0, 1, 2
SF_FLUID, SF_SOLID = 0, 1
STAGES = {"forces": 0, "ibm": 1, "lbm": 2, "scalar": 3, "move": 4}

# every entry point the header declares (checked by tests/test_capi_cpu.py)
EXPORTS = (
    "sf_last_error", "sf_device_count", "sf_create", "sf_destroy", "sf_set_flags", "sf_set_obstacles",
    "sf_set_populations", "sf_get_populations", "sf_get_macroscopic", "sf_set_macroscopic",
    "sf_set_capsules", "sf_get_nodes", "sf_set_nodes", "sf_get_membrane_state",
    "sf_get_force_state", "sf_set_force_state",
    "sf_set_scalar", "sf_get_scalar", "sf_step", "sf_set_graph_mode", "sf_synchronize", "sf_graph_captures",
    "sf_stage", "sf_last_capped", "sf_field_stats", "sf_field_stats_noalloc", "sf_global_stats", "sf_step_count", "sf_kernel_times",
    "sf_host_alloc", "sf_host_free", "sf_set_probes", "sf_snapshot_async", "sf_snapshot_wait",
    "sf_set_timing", "sf_autotune", "sf_set_lbm_launch", "sf_set_lbm_pairs", "sf_set_capsule_launch", "sf_probe_math", "sf_copy_ceiling",
    "sf_mdf_set_control", "sf_mdf_set_tail", "sf_mdf_stats", "sf_mdf_collect", "sf_mdf_samples", "sf_mdf_set_model",
    "sf_mlp_create", "sf_mlp_destroy", "sf_mlp_num_params", "sf_mlp_get_params", "sf_mlp_set_params",
    "sf_mlp_predict", "sf_mlp_loss_grad", "sf_mlp_train",
    "sf_batch_create", "sf_batch_destroy", "sf_batch_step", "sf_batch_synchronize",
    "sf_batch_set_graph_mode", "sf_batch_autotune", "sf_batch_set_lbm_launch", "sf_batch_set_lbm_flat",
    "sf_batch_info",
)


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("dt", C.c_double), ("ibm_iterations", C.c_int),
        ("tau", C.c_double), ("rho0", C.c_double), ("boundary_type", C.c_int),
        ("body_force_x", C.c_double), ("body_force_y", C.c_double),
        ("collision_model", C.c_int),
        ("mrt_se", C.c_double), ("mrt_s_eps", C.c_double), ("mrt_sq", C.c_double),
        ("periodic_y", C.c_int), ("max_lattice_force", C.c_double),
        ("rep_epsilon", C.c_double), ("rep_sigma", C.c_double), ("rep_r_cut", C.c_double),
        ("rep_power", C.c_int), ("rep_damping_normal", C.c_double),
        ("rep_friction_coeff", C.c_double),
        ("scalar_enabled", C.c_int), ("n_species", C.c_int),
        ("diffusivity", C.c_double * 4), ("inlet_concentration", C.c_double * 4),
        ("scalar_periodic_y", C.c_int), ("precision", C.c_int),
        ("top_wall_velocity", C.c_double), ("bottom_wall_velocity", C.c_double),
        ("inlet_velocity", C.c_double), ("outlet_density", C.c_double),
        ("use_interpolated_bb", C.c_int),
        ("capsule_periodic_x", C.c_int),
    ]


class CapsuleDesc(C.Structure):
    _fields_ = [
        ("model", C.c_int), ("use_helfrich", C.c_int), ("is_rigid", C.c_int), ("type", C.c_int),
        ("k_stretch", C.c_double), ("G_s", C.c_double), ("C_skalak", C.c_double),
        ("k_bend", C.c_double), ("k_area", C.c_double), ("k_perimeter", C.c_double),
        ("gamma_visc", C.c_double), ("eta_membrane", C.c_double),
        ("wlc_L_max_ratio", C.c_double), ("wlc_kBT_p", C.c_double), ("wlc_k_pow", C.c_double),
        ("rest_area", C.c_double), ("rest_perimeter", C.c_double),
    ]


_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int32)
_lib = None


def load():
    """Load libsoftflow_hip.so and declare the prototypes."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -m softflow_amd.build`")
    L = C.CDLL(LIB_PATH)
    L.sf_last_error.restype = C.c_char_p
    proto = {
        "sf_device_count": [C.POINTER(C.c_int)],
        "sf_create": [C.POINTER(C.c_void_p), C.c_int, C.POINTER(Params)],
        "sf_destroy": [C.c_void_p],
        "sf_set_flags": [C.c_void_p, _IP],
        "sf_set_obstacles": [C.c_void_p, C.c_int, _IP, _IP, _DP],
        "sf_set_populations": [C.c_void_p, _DP],
        "sf_get_populations": [C.c_void_p, _DP],
        "sf_get_macroscopic": [C.c_void_p, _DP, _DP, _DP, _DP, _DP],
        "sf_set_macroscopic": [C.c_void_p, _DP, _DP, _DP],
        "sf_set_capsules": [C.c_void_p, C.c_int, _IP, C.POINTER(CapsuleDesc), _DP, _DP, _DP, _DP,
                            _DP, _DP, _DP, _DP, _IP],
        "sf_get_nodes": [C.c_void_p, _DP, _DP, _DP, _DP, _DP, _DP],
        "sf_set_nodes": [C.c_void_p, _DP, _DP, _DP, _DP],
        "sf_get_membrane_state": [C.c_void_p, _DP, _IP],
        "sf_get_force_state": [C.c_void_p, _DP],
        "sf_set_force_state": [C.c_void_p, _DP],
        "sf_set_scalar": [C.c_void_p, C.c_int, _DP, _DP],
        "sf_get_scalar": [C.c_void_p, C.c_int, _DP, _DP],
        "sf_step": [C.c_void_p, C.c_int],
        "sf_set_graph_mode": [C.c_void_p, C.c_int],
        "sf_synchronize": [C.c_void_p],
        "sf_stage": [C.c_void_p, C.c_int],
        "sf_last_capped": [C.c_void_p, _IP],
        "sf_field_stats": [C.c_void_p, _DP],
        "sf_field_stats_noalloc": [C.c_void_p, _DP],
        "sf_global_stats": [C.c_void_p, C.c_double, _DP],
        "sf_step_count": [C.c_void_p, C.POINTER(C.c_int64)],
        "sf_graph_captures": [C.c_void_p, C.POINTER(C.c_int64)],
        "sf_host_alloc": [C.POINTER(C.c_void_p), C.c_uint64],
        "sf_host_free": [C.c_void_p],
        "sf_set_probes": [C.c_void_p, C.c_int, _IP],
        "sf_snapshot_async": [C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)],
        "sf_snapshot_wait": [C.c_void_p, C.c_int64],
        "sf_kernel_times": [C.c_void_p, C.c_int, C.POINTER(C.c_char_p), _DP, C.POINTER(C.c_int)],
        "sf_set_timing": [C.c_void_p, C.c_int],
        "sf_autotune": [C.c_void_p, C.c_int, C.c_char_p, C.c_int],
        "sf_set_lbm_launch": [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int],
        "sf_set_lbm_pairs": [C.c_void_p, C.c_int],
        "sf_set_capsule_launch": [C.c_void_p, C.c_int, C.c_int],
        "sf_probe_math": [C.c_int, C.c_int, _DP, _DP, _DP],
        "sf_copy_ceiling": [C.c_int, C.c_uint64, C.c_int, C.POINTER(C.c_double)],
        "sf_mdf_set_control": [C.c_void_p, C.c_double, C.c_int, C.c_int],
        "sf_mdf_set_tail": [C.c_void_p, C.c_int],
        "sf_mdf_stats": [C.c_void_p, _DP],
        "sf_mdf_collect": [C.c_void_p, C.c_int64],
        "sf_mdf_samples": [C.c_void_p, C.POINTER(C.c_int64), _FP],
        "sf_mdf_set_model": [C.c_void_p, C.c_void_p, C.c_int],
        "sf_mlp_create": [C.POINTER(C.c_void_p), C.c_int, C.c_uint64],
        "sf_mlp_destroy": [C.c_void_p],
        "sf_mlp_num_params": [C.POINTER(C.c_int), C.POINTER(C.c_int)],
        "sf_mlp_get_params": [C.c_void_p, _FP, _FP],
        "sf_mlp_set_params": [C.c_void_p, _FP, _FP],
        "sf_mlp_predict": [C.c_void_p, C.c_int, _FP, _FP],
        "sf_mlp_loss_grad": [C.c_void_p, C.c_int, _FP, _FP, _DP, _FP],
        "sf_mlp_train": [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_uint64, _DP],
        "sf_batch_create": [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int],
        "sf_batch_destroy": [C.c_void_p],
        "sf_batch_step": [C.c_void_p, C.c_int],
        "sf_batch_synchronize": [C.c_void_p],
        "sf_batch_set_graph_mode": [C.c_void_p, C.c_int],
        "sf_batch_autotune": [C.c_void_p, C.c_int, C.c_char_p, C.c_int],
        "sf_batch_set_lbm_launch": [C.c_void_p, C.c_int, C.c_int, C.c_int],
        "sf_batch_set_lbm_flat": [C.c_void_p, C.c_int],
        "sf_batch_info": [C.c_void_p, C.POINTER(C.c_int64)],
    }
    for name, args in proto.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = L
    return L


class SoftFlowError(RuntimeError):
    pass


def _check(rc: int, what: str):
    if rc != SF_OK:
        raise SoftFlowError(f"{what} failed ({rc}): {load().sf_last_error().decode()}")


def _dp(a):
    return None if a is None else a.ctypes.data_as(_DP)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def device_count() -> int:
    n = C.c_int(0)
    rc = load().sf_device_count(C.byref(n))
    return n.value if rc == SF_OK else 0


def make_params(**kw) -> Params:
    """sf_params with the reference's defaults (core/parameters.h)."""
    p = Params()
    p.nx, p.ny, p.dt, p.ibm_iterations = 200, 50, 1.0, 1
    p.tau, p.rho0, p.boundary_type = 0.8, 1.0, SF_PERIODIC
    p.mrt_se = p.mrt_s_eps = p.mrt_sq = 1.4
    p.max_lattice_force = 0.01
    p.rep_epsilon, p.rep_sigma, p.rep_r_cut, p.rep_power = 0.05, 1.0, 3.5, 4
    p.n_species = 1
    p.diffusivity[0] = 0.01
    p.precision = SF_FP64
    for k, v in kw.items():
        if k in ("diffusivity", "inlet_concentration"):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(p, k, v)
    return p


class Context:
    """One device context (one Simulation's state) driven through the C-ABI."""

    def __init__(self, params: Params, device: int = 0):
        self.L = load()
        self.params = params
        self.nx, self.ny = params.nx, params.ny
        self.N = params.nx * params.ny
        h = C.c_void_p()
        _check(self.L.sf_create(C.byref(h), device, C.byref(params)), "sf_create")
        self.h = h
        self.K = 0
        self.ncaps = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.sf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_flags(self, flags):
        a = np.ascontiguousarray(flags, dtype=np.int32).reshape(-1)
        _check(self.L.sf_set_flags(self.h, a.ctypes.data_as(_IP)), "sf_set_flags")

    def set_obstacles(self, obstacles):
        """obstacles: [("circle", cx, cy, r) | ("rect", x0, y0, x1, y1) | ("polygon", [(x, y), ...])]"""
        kinds, offs, coords = [], [0], []
        for o in obstacles:
            kind = {"circle": 0, "rect": 1, "polygon": 2}[o[0]]
            vals = [float(v) for xy in o[1] for v in xy] if kind == 2 else [float(v) for v in o[1:]]
            kinds.append(kind)
            coords.extend(vals)
            offs.append(len(coords))
        k = np.asarray(kinds, dtype=np.int32)
        f = np.asarray(offs, dtype=np.int32)
        c = np.asarray(coords if coords else [0.0], dtype=np.float64)
        _check(self.L.sf_set_obstacles(self.h, len(kinds), k.ctypes.data_as(_IP), f.ctypes.data_as(_IP),
                                       _dp(c)), "sf_set_obstacles")

    def set_populations(self, f):
        a = _f64(f).reshape(-1)
        _check(self.L.sf_set_populations(self.h, _dp(a)), "sf_set_populations")

    def populations(self):
        a = np.empty(9 * self.N)
        _check(self.L.sf_get_populations(self.h, _dp(a)), "sf_get_populations")
        return a.reshape(9, self.ny, self.nx)

    def macroscopic(self):
        arrs = [np.empty(self.N) for _ in range(5)]
        _check(self.L.sf_get_macroscopic(self.h, *[_dp(a) for a in arrs]), "sf_get_macroscopic")
        return {k: a.reshape(self.ny, self.nx) for k, a in zip(("rho", "ux", "uy", "fx", "fy"), arrs)}

    def set_macroscopic(self, rho, ux, uy):
        arrs = [_f64(a).reshape(-1) for a in (rho, ux, uy)]
        _check(self.L.sf_set_macroscopic(self.h, *[_dp(a) for a in arrs]), "sf_set_macroscopic")

    def set_capsules(self, off, desc, x, y, vx, vy, rest_len, rest_angle=None, rest_ds=None,
                     prev_len=None, has_prev=None):
        off = np.ascontiguousarray(off, dtype=np.int32)
        ncaps = len(off) - 1
        K = int(off[-1]) if ncaps > 0 else 0
        D = (CapsuleDesc * max(ncaps, 1))(*desc)
        arrs = [_f64(a) for a in (x, y, vx, vy, rest_len)]
        ra = _f64(rest_angle) if rest_angle is not None else np.zeros(K)
        rd = _f64(rest_ds) if rest_ds is not None else np.zeros(K)
        pl = _f64(prev_len) if prev_len is not None else arrs[4].copy()
        hp = np.ascontiguousarray(has_prev if has_prev is not None else np.zeros(ncaps), dtype=np.int32)
        _check(self.L.sf_set_capsules(self.h, ncaps, off.ctypes.data_as(_IP), D, *[_dp(a) for a in arrs],
                                      _dp(ra), _dp(rd), _dp(pl), hp.ctypes.data_as(_IP)),
               "sf_set_capsules")
        self.K, self.ncaps = K, ncaps

    def nodes(self):
        arrs = [np.empty(self.K) for _ in range(6)]
        _check(self.L.sf_get_nodes(self.h, *[_dp(a) for a in arrs]), "sf_get_nodes")
        return dict(zip(("x", "y", "vx", "vy", "fx", "fy"), arrs))

    def set_nodes(self, x=None, y=None, vx=None, vy=None):
        arrs = [None if a is None else _f64(a) for a in (x, y, vx, vy)]
        _check(self.L.sf_set_nodes(self.h, *[_dp(a) for a in arrs]), "sf_set_nodes")

    def set_scalar(self, species, g, Cc):
        g = _f64(g).reshape(-1)
        Cc = _f64(Cc).reshape(-1)
        _check(self.L.sf_set_scalar(self.h, species, _dp(g), _dp(Cc)), "sf_set_scalar")

    def scalar(self, species=0, with_g=False):
        Cc = np.empty(self.N)
        g = np.empty(9 * self.N) if with_g else None
        _check(self.L.sf_get_scalar(self.h, species, _dp(Cc), _dp(g)), "sf_get_scalar")
        Cc = Cc.reshape(self.ny, self.nx)
        return (Cc, g) if with_g else Cc

    def step(self, n=1):
        _check(self.L.sf_step(self.h, n), "sf_step")

    def stage(self, name):
        _check(self.L.sf_stage(self.h, STAGES[name]), "sf_stage")

    def set_graph_mode(self, on: bool):
        _check(self.L.sf_set_graph_mode(self.h, 1 if on else 0), "sf_set_graph_mode")

    def synchronize(self):
        _check(self.L.sf_synchronize(self.h), "sf_synchronize")

    def last_capped(self):
        c = C.c_int32(0)
        _check(self.L.sf_last_capped(self.h, C.byref(c)), "sf_last_capped")
        return c.value

    def field_stats(self):
        out = np.empty(5)
        _check(self.L.sf_field_stats(self.h, _dp(out)), "sf_field_stats")
        return out

    def field_stats_noalloc(self):
        out = np.empty(5)
        _check(self.L.sf_field_stats_noalloc(self.h, _dp(out)), "sf_field_stats_noalloc")
        return out

    def set_timing(self, on: bool):
        _check(self.L.sf_set_timing(self.h, 1 if on else 0), "sf_set_timing")

    def autotune(self, reps=10):
        buf = C.create_string_buffer(4096)
        _check(self.L.sf_autotune(self.h, reps, buf, 4096), "sf_autotune")
        return buf.value.decode()

    def kernel_times(self):
        names = (C.c_char_p * 32)()
        ms = np.zeros(32)
        n = C.c_int(0)
        _check(self.L.sf_kernel_times(self.h, 32, names, _dp(ms), C.byref(n)), "sf_kernel_times")
        return {names[i].decode(): float(ms[i]) for i in range(min(n.value, 32))}


class Batch:
    """A batched ensemble of Contexts (sf_batch_*); the contexts must outlive it."""

    def __init__(self, contexts):
        self.L = load()
        self.members = list(contexts)
        arr = (C.c_void_p * len(self.members))(*[c.h.value for c in self.members])
        h = C.c_void_p()
        _check(self.L.sf_batch_create(C.byref(h), arr, len(self.members)), "sf_batch_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.sf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        _check(self.L.sf_batch_step(self.h, n), "sf_batch_step")

    def synchronize(self):
        _check(self.L.sf_batch_synchronize(self.h), "sf_batch_synchronize")

    def set_graph_mode(self, on: bool):
        _check(self.L.sf_batch_set_graph_mode(self.h, 1 if on else 0), "sf_batch_set_graph_mode")

    def set_lbm_launch(self, bx, by, xcd_banded=True):
        _check(self.L.sf_batch_set_lbm_launch(self.h, bx, by, 1 if xcd_banded else 0),
               "sf_batch_set_lbm_launch")

    def autotune(self, reps=10):
        buf = C.create_string_buffer(4096)
        _check(self.L.sf_batch_autotune(self.h, reps, buf, 4096), "sf_batch_autotune")
        return buf.value.decode()

    def info(self):
        out = (C.c_int64 * 4)()
        _check(self.L.sf_batch_info(self.h, out), "sf_batch_info")
        return list(out)


def probe_math(op: int, a, b=None):
    a = _f64(a)
    bb = _f64(b) if b is not None else None
    out = np.empty_like(a)
    _check(load().sf_probe_math(op, a.size, _dp(a), _dp(bb), _dp(out)), "sf_probe_math")
    return out


def copy_ceiling(nbytes: int = 1 << 30, reps: int = 10, device: int = 0) -> float:
    """Achievable HBM bandwidth (GB/s) of a nontemporal 16-byte streaming copy."""
    out = C.c_double()
    _check(load().sf_copy_ceiling(device, nbytes, reps, C.byref(out)), "sf_copy_ceiling")
    return out.value
