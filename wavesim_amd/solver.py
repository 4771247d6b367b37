"""Handle-level wrapper of the HIP solver (wsim_solver_* in include/wavesim_mbs.h).

``MBSSolver`` owns the device buffers of one problem (gamma, wiggle tables,
E', dE', the work field and the scaled sources of ``nrhs`` right-hand sides)
and runs the MBS loop of reference WaveSimBase.m:157-197 on the GPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import capi


@dataclass
class RhsResult:
    iterations: int
    converged: bool
    diff_energy: np.ndarray
    energy_every: np.ndarray
    diverged: bool = False  # stopped by the failure exit of the stop rule


class MBSSolver:
    """dtype: np.complex64 (the reference's default, single_precision=true)
    or np.complex128 (single_precision=false, Simulation.m:31, 221-252): the
    fields, gamma, the wiggle tables and the sources are held and iterated in
    that precision on the GPU."""

    def __init__(self, N: Sequence[int], gamma, wiggles, epsilon: float, k02e: complex,
                 roi, energy_interval: int = 8, nrhs: int = 1, dtype=np.complex64):
        L = capi.lib()
        N = [int(x) for x in N]
        if len(N) == 3:
            N = N + [1]
        self.N = N
        self.nrhs = int(nrhs)
        self.dtype = np.dtype(dtype).type
        if self.dtype not in (np.complex64, np.complex128):
            raise ValueError("dtype must be complex64 or complex128")
        self.f64 = self.dtype is np.complex128
        g = capi.c64(np.asarray(gamma).reshape(N[:3]), self.dtype)
        self._w = capi.Wiggles(wiggles, self.dtype)
        d = capi.desc_f64_t() if self.f64 else capi.desc_t()
        d.N = (C.c_int64 * 4)(*N)
        d.nrhs = self.nrhs
        d.nwiggles = self._w.n
        d.epsilon = float(epsilon)
        ctype = capi.cdouble if self.f64 else capi.cfloat
        d.k02e = ctype(float(np.real(k02e)), float(np.imag(k02e)))
        roi = np.asarray(roi, dtype=np.int64)
        for r in range(2):
            for c in range(4):
                d.roi[r][c] = int(roi[r, c])
        d.energy_interval = int(energy_interval)
        h = C.c_void_p()
        create = L.wsim_solver_create_f64 if self.f64 else L.wsim_solver_create
        capi.check(create(C.byref(h), C.byref(d), capi.ptr(g), self._w.arr))
        self._h = h
        self._lib = L
        self.nw = self._w.n

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.wsim_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- setup --------------------------------------------------------------
    def set_source(self, rhs: int, positions, values):
        """positions: 1-based [y,x,z,pol] grid positions; values: arrays in
        [y,x,z,pol] index order (already shifted and cropped to the grid)."""
        p = capi.Patches(positions, values, self.dtype)
        fn = self._lib.wsim_solver_set_source_f64 if self.f64 else self._lib.wsim_solver_set_source
        capi.check(fn(self._h, int(rhs), p.n, p.arr))

    # -- run ------------------------------------------------------------------
    def solve(self, max_iter: int, energy_threshold: float, batch: int = 0,
              callback: Optional[Callable] = None, callback_interval: float = 0.0,
              callback_type: str = "NoCallback"):
        if callback is not None:
            def _cb(typ, it, de, n, thr, user):
                hist = np.ctypeslib.as_array(de, shape=(n,)).copy() if n > 0 else np.zeros(0)
                return 1 if callback(int(it), hist) else 0
            cb = capi.CALLBACK(_cb)
        else:
            cb = capi.NULL_CALLBACK
        rc = self._lib.wsim_solver_solve(self._h, int(max_iter), float(energy_threshold), int(batch),
                                         callback_type.encode(), float(callback_interval), cb, None)
        if rc == capi.WSIM_ERR_ABORTED:
            return False
        capi.check(rc)
        return True

    def result(self, rhs: int = 0, cap: int = 1 << 22) -> RhsResult:
        n = C.c_int64()
        cv = C.c_int()
        capi.check(self._lib.wsim_solver_result(self._h, int(rhs), C.byref(n), C.byref(cv), None, None, 0))
        m = int(min(n.value, cap))
        de = np.zeros(max(m, 1))
        ea = np.zeros(max(m, 1))
        capi.check(self._lib.wsim_solver_result(
            self._h, int(rhs), C.byref(n), C.byref(cv),
            de.ctypes.data_as(C.POINTER(C.c_double)), ea.ctypes.data_as(C.POINTER(C.c_double)), m))
        return RhsResult(int(n.value), cv.value == 1, de[:m], ea[:m], cv.value < 0)

    def field(self, rhs: int = 0) -> np.ndarray:
        """E' (primed, full grid) as a [y,x,z,pol] array of the solver's dtype."""
        out = np.empty(self.N, dtype=self.dtype, order="F")
        fn = self._lib.wsim_solver_get_field_f64 if self.f64 else self._lib.wsim_solver_get_field
        capi.check(fn(self._h, int(rhs), capi.ptr(out)))
        return out

    def device_ptr(self, which: int, rhs: int = 0) -> int:
        p = C.c_void_p()
        capi.check(self._lib.wsim_solver_device_ptr(self._h, int(which), int(rhs), C.byref(p)))
        return int(p.value)

    def propagate(self, wiggle: int, E) -> np.ndarray:
        """One application of the propagation operator (WaveSim.m:15-36 /
        WaveSimVector.m:19-51) of wiggle `wiggle` (0-based)."""
        x = capi.c64(np.asarray(E).reshape(self.N), self.dtype)
        out = np.empty(self.N, dtype=self.dtype, order="F")
        fn = self._lib.wsim_solver_propagate_f64 if self.f64 else self._lib.wsim_solver_propagate
        capi.check(fn(self._h, int(wiggle), capi.ptr(x), capi.ptr(out)))
        return out

    def set_medium(self, gamma, wiggles, epsilon: float, k02e: complex):
        """Replace the medium (same grid and wiggle count); sources are cleared."""
        g = capi.c64(np.asarray(gamma).reshape(self.N[:3]))
        w = capi.Wiggles(wiggles)
        capi.check(self._lib.wsim_solver_set_medium(self._h, capi.ptr(g), float(epsilon),
                                                    capi.cfloat(float(np.real(k02e)), float(np.imag(k02e))),
                                                    w.arr))
        self._w = w

    # -- warm start -------------------------------------------------------------
    def set_warm_start(self, rhs: int, E0=None, device_ptr: Optional[int] = None):
        """Start rhs from the unprimed field E0 ([y,x,z,pol] host array) or
        from a device buffer holding it in the solver's layout."""
        if device_ptr is not None:
            capi.check(self._lib.wsim_solver_set_warm_start(self._h, int(rhs), C.c_void_p(device_ptr), 1))
        else:
            x = capi.c64(np.asarray(E0).reshape(self.N))
            capi.check(self._lib.wsim_solver_set_warm_start(self._h, int(rhs), capi.ptr(x), 0))

    def clear_warm_start(self):
        capi.check(self._lib.wsim_solver_clear_warm_start(self._h))

    def set_reference_energy(self, rhs: int, energy: float):
        capi.check(self._lib.wsim_solver_set_reference_energy(self._h, int(rhs), float(energy)))

    # -- FFT plans ---------------------------------------------------------------
    def tune(self, reps: int = 4):
        """Autotune the per-axis radix order and tile width on this problem;
        returns (plan text, ms per iteration)."""
        buf = C.create_string_buffer(512)
        ms = C.c_double()
        capi.check(self._lib.wsim_solver_tune(self._h, int(reps), buf, 512, C.byref(ms)))
        return buf.value.decode(), ms.value

    @property
    def plan(self) -> str:
        buf = C.create_string_buffer(512)
        capi.check(self._lib.wsim_solver_get_plan(self._h, buf, 512))
        return buf.value.decode()

    @plan.setter
    def plan(self, text: str):
        capi.check(self._lib.wsim_solver_set_plan(self._h, text.encode()))

    @property
    def fixed_launches(self) -> int:
        """Pass launches of one iteration that run plan-specialised kernels."""
        n = self._lib.wsim_solver_fixed_launches(self._h)
        if n < 0:
            capi.check(n)
        return n

    def profile(self, reps: int = 10):
        """[(kernel name, mean ms, algorithmic bytes)] for one iteration."""
        K = 16
        ms = (C.c_double * K)()
        by = (C.c_double * K)()
        nm = (C.c_char_p * K)()
        nk = C.c_int()
        capi.check(self._lib.wsim_solver_profile(self._h, int(reps), K, ms, by, nm, C.byref(nk)))
        return [(nm[i].decode(), ms[i], by[i]) for i in range(min(nk.value, K))]


class SolverGroup:
    """Independent problems (MBSSolver objects of any grid sizes) advanced
    together: their passes share launches and one control launch serves all
    of them (wsim_group_* in include/wavesim_mbs.h).  After ``solve`` every
    solver's ``result``/``field`` read exactly as after its own ``solve``.
    The reference runs several problems as a loop of exec calls
    (examples/pulse2d.m:31-40)."""

    def __init__(self, solvers: Sequence[MBSSolver]):
        L = capi.lib()
        self.solvers = list(solvers)
        hs = (C.c_void_p * len(self.solvers))(*[s._h for s in self.solvers])
        h = C.c_void_p()
        capi.check(L.wsim_group_create(C.byref(h), hs, len(self.solvers)))
        self._h = h
        self._lib = L

    def solve(self, max_iter, energy_threshold, batch: int = 0):
        """max_iter, energy_threshold: one value for all problems or one per problem."""
        if not getattr(self, "_h", None):
            raise RuntimeError("the group is closed")
        if any(not getattr(s, "_h", None) for s in self.solvers):
            raise RuntimeError("a solver of this group was closed; close the group before its solvers")
        n = len(self.solvers)
        mi = np.broadcast_to(np.asarray(max_iter, dtype=np.int64), (n,)).copy()
        th = np.broadcast_to(np.asarray(energy_threshold, dtype=np.float64), (n,)).copy()
        capi.check(self._lib.wsim_group_solve(self._h, mi.ctypes.data_as(C.POINTER(C.c_int64)),
                                              th.ctypes.data_as(C.POINTER(C.c_double)), int(batch)))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.wsim_group_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
