"""ctypes binding of ``libwavesim_mbs.so`` (declared in include/wavesim_mbs.h).

The library is required: importing this module raises if it is missing, and
there is no CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("WSIM_LIB", os.path.join(_HERE, "libwavesim_mbs.so"))

WSIM_OK = 0
WSIM_ERR_UNSUPPORTED = -3
ERRORS = {-1: "WSIM_ERR_ARG", -2: "WSIM_ERR_HIP", -3: "WSIM_ERR_UNSUPPORTED", -4: "WSIM_ERR_NOMEM",
          -5: "WSIM_ERR_ABORTED", -6: "WSIM_ERR_DIVERGED"}
WSIM_ERR_ABORTED = -5  # the callback asked to stop
WSIM_ERR_DIVERGED = -6  # the stop rule's failure exit (non-finite or growing increment energy)


class WsimError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class cfloat(C.Structure):
    _fields_ = [("re", C.c_float), ("im", C.c_float)]


# (also the layout of wsim_wiggle_f64: six pointers)
class wiggle_t(C.Structure):
    _fields_ = [("pxe", C.c_void_p), ("pye", C.c_void_p), ("pze", C.c_void_p),
                ("gx", C.c_void_p), ("gy", C.c_void_p), ("gz", C.c_void_p)]


# (also the layout of wsim_source_patch_f64)
class patch_t(C.Structure):
    _fields_ = [("pos", C.c_double * 4), ("size", C.c_int64 * 4), ("values", C.c_void_p)]


class desc_t(C.Structure):
    _fields_ = [("N", C.c_int64 * 4), ("nrhs", C.c_int), ("nwiggles", C.c_int),
                ("epsilon", C.c_float), ("k02e", cfloat), ("roi", (C.c_int64 * 4) * 2),
                ("energy_interval", C.c_int)]


class cdouble(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class desc_f64_t(C.Structure):
    _fields_ = [("N", C.c_int64 * 4), ("nrhs", C.c_int), ("nwiggles", C.c_int),
                ("epsilon", C.c_double), ("k02e", cdouble), ("roi", (C.c_int64 * 4) * 2),
                ("energy_interval", C.c_int)]


class problem_desc_t(C.Structure):
    _fields_ = [("shape", C.c_int64 * 3), ("pixel_size", C.c_double), ("lambda_", C.c_double),
                ("boundary_widths", C.c_int64 * 3), ("epsilon", C.c_double), ("epsilonmin", C.c_double),
                ("acc", C.c_int * 3), ("vector", C.c_int), ("has_roi", C.c_int), ("roi", (C.c_int64 * 4) * 2)]


class problem_info_t(C.Structure):
    _fields_ = [("N", C.c_int64 * 4), ("roi", (C.c_int64 * 4) * 2), ("Bl", C.c_int64 * 3), ("Br", C.c_int64 * 3),
                ("nwiggles", C.c_int), ("epsilon", C.c_float), ("k02e", cfloat), ("k0c", C.c_double),
                ("iterations_per_cycle", C.c_double)]


class pstd_desc_t(C.Structure):
    _fields_ = [("N", C.c_int64 * 3), ("roi", (C.c_int64 * 4) * 2), ("iterations_per_cycle", C.c_double),
                ("energy_interval", C.c_int)]


CALLBACK = C.CFUNCTYPE(C.c_int, C.c_char_p, C.c_int64, C.POINTER(C.c_double), C.c_int64,
                       C.c_double, C.c_void_p)

SYMBOLS = {
    "wsim_run_algorithm": (C.c_int, [C.POINTER(C.c_double), C.c_void_p, C.c_int, C.POINTER(wiggle_t),
                                     C.c_int, C.POINTER(patch_t), C.c_float, cfloat,
                                     C.POINTER(C.c_double), C.c_double, C.c_double, C.c_char_p,
                                     C.c_double, CALLBACK, C.c_void_p, C.c_void_p,
                                     C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_int64)]),
    "wsim_run_algorithm_noalloc": (C.c_int, [C.POINTER(C.c_double), C.c_void_p, C.c_int, C.POINTER(wiggle_t),
                                             C.c_int, C.POINTER(patch_t), C.c_float, cfloat,
                                             C.POINTER(C.c_double), C.c_double, C.c_double, C.c_char_p,
                                             C.c_double, CALLBACK, C.c_void_p, C.c_void_p,
                                             C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_int64),
                                             C.c_void_p]),
    "wsim_solver_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(desc_t), C.c_void_p,
                                     C.POINTER(wiggle_t)]),
    "wsim_solver_set_source": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(patch_t)]),
    "wsim_solver_create_f64": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(desc_f64_t), C.c_void_p,
                                         C.POINTER(wiggle_t)]),
    "wsim_solver_set_source_f64": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(patch_t)]),
    "wsim_solver_get_field_f64": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "wsim_solver_propagate_f64": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "wsim_solver_element_size": (C.c_int, [C.c_void_p]),
    "wsim_solver_fixed_launches": (C.c_int, [C.c_void_p]),
    "wsim_solver_solve": (C.c_int, [C.c_void_p, C.c_int64, C.c_double, C.c_int64, C.c_char_p,
                                    C.c_double, CALLBACK, C.c_void_p]),
    "wsim_solver_result": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int),
                                     C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64]),
    "wsim_solver_get_field": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "wsim_solver_device_ptr": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "wsim_solver_propagate": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "wsim_solver_profile": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.POINTER(C.c_char_p), C.POINTER(C.c_int)]),
    "wsim_solver_tune": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_double)]),
    "wsim_solver_get_plan": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "wsim_solver_set_plan": (C.c_int, [C.c_void_p, C.c_char_p]),
    "wsim_solver_set_medium": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, cfloat, C.POINTER(wiggle_t)]),
    "wsim_solver_set_warm_start": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "wsim_solver_clear_warm_start": (C.c_int, [C.c_void_p]),
    "wsim_solver_set_reference_energy": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "wsim_solver_destroy": (C.c_int, [C.c_void_p]),
    "wsim_group_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int]),
    "wsim_group_solve": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_double), C.c_int64]),
    "wsim_group_destroy": (C.c_int, [C.c_void_p]),
    "wsim_problem_query": (C.c_int, [C.POINTER(problem_desc_t), C.POINTER(problem_info_t)]),
    "wsim_problem_build": (C.c_int, [C.POINTER(problem_desc_t), C.c_void_p, C.POINTER(problem_info_t),
                                     C.c_void_p, C.POINTER(wiggle_t)]),
    "wsim_pstd_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(pstd_desc_t)] + [C.c_void_p] * 6),
    "wsim_pstd_set_source": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(patch_t)]),
    "wsim_pstd_run": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "wsim_pstd_result": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.c_int64]),
    "wsim_pstd_destroy": (C.c_int, [C.c_void_p]),
    "wsim_pstd_fixed_launches": (C.c_int, [C.c_void_p]),
    "wsim_last_error": (C.c_char_p, []),
    "wsim_abi_version": (C.c_int, []),
    "wsim_selftest": (C.c_int, []),
}

_lib = None


def lib():
    """Load the solver library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m wavesim_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SYMBOLS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> int:
    if rc != WSIM_OK:
        raise WsimError(rc, lib().wsim_last_error().decode(errors="replace"))
    return rc


def c64(a, dtype=np.complex64) -> np.ndarray:
    """complex64 (or ``dtype``), MATLAB column-major (Fortran) order."""
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def f32(a, dtype=np.float32) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=dtype).ravel())


def real_dtype(cdtype) -> type:
    return np.float64 if np.dtype(cdtype) == np.complex128 else np.float32


def ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class Wiggles:
    """Keeps the host tables of the wiggle descriptors alive for a call."""

    def __init__(self, wiggles, dtype=np.complex64):
        self.keep = []
        arr = (wiggle_t * len(wiggles))()
        rdt = real_dtype(dtype)
        for i, w in enumerate(wiggles):
            t = [f32(w["pxe"], rdt), f32(w["pye"], rdt), f32(w["pze"], rdt),
                 np.ascontiguousarray(np.asarray(w["gx"], dtype).ravel()),
                 np.ascontiguousarray(np.asarray(w["gy"], dtype).ravel()),
                 np.ascontiguousarray(np.asarray(w["gz"], dtype).ravel())]
            self.keep += t
            arr[i] = wiggle_t(*[ptr(x) for x in t])
        self.arr = arr
        self.n = len(wiggles)


class Patches:
    """Source patches: list of (1-based position[4], value array in [y,x,z,pol] order)."""

    def __init__(self, positions, values, dtype=np.complex64):
        self.keep = []
        n = len(positions)
        arr = (patch_t * max(n, 1))()
        for i, (p, v) in enumerate(zip(positions, values)):
            v = np.asarray(v)
            v4 = v.reshape(list(v.shape) + [1] * (4 - v.ndim)) if v.ndim < 4 else v
            vf = c64(v4, dtype)
            self.keep.append(vf)
            arr[i].pos = (C.c_double * 4)(*[float(x) for x in p])
            arr[i].size = (C.c_int64 * 4)(*[int(x) for x in v4.shape])
            arr[i].values = ptr(vf)
        self.arr = arr
        self.n = n


NULL_CALLBACK = CALLBACK()
