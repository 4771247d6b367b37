"""Host-side mirror of the reference's PSTD comparison solver (PSTD.m), running
its time stepping on the GPU through ``include/wavesim_pstd.h``.

The constructor computes what PSTD.m:22-77 computes -- the padded medium, the
time step dt (95% of the stability limit), the coefficient maps c1, c2, c3 and
the spectral Laplacian's coordinates -- and ``exec`` runs
``PSTD.run_algorithm`` (PSTD.m:79-127) on the HIP axis-pass kernels: each step
is ifftn(-|p|^2 fftn(E)) plus the ramped CW source, the leapfrog update and the
step energy, until Simulation.next's stopping rule with PSTD's
can_terminate (real part of the source amplitude above 0.9) ends the run.
Conventions and option names are the reference's.  Only the default source
amplitude (PSTD.default_source) is supported.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import capi
from .simulation import Medium, Source, _make4


@dataclass
class PSTDState:
    it: int = 0
    iterations: int = 0
    converged: bool = False
    diff_energy: np.ndarray = field(default_factory=lambda: np.zeros(0))
    time: float = 0.0
    run_time: float = 0.0   # the time stepping alone (device-resident), seconds
    E_full: Optional[np.ndarray] = None


def default_source(ncycle: float) -> complex:
    """PSTD.default_source (PSTD.m:130-137)."""
    A = complex(math.cos(-2 * math.pi * ncycle), math.sin(-2 * math.pi * ncycle))
    return A * (0.5 - 0.5 * math.cos(math.pi * min(ncycle / 100.0, 1.0)))


class PSTD:
    """PSTD(refractive_index, options) (PSTD.m:22-77)."""

    def __init__(self, refractive_index, options: dict):
        opt = dict(options)
        # PSTD.m:32-34
        if "boundary_type" not in opt or "boundary_strength" in opt:
            raise ValueError("options field boundary_type and boundary_strength needs to be specified")
        opt["ACC"] = False                                               # PSTD.m:35
        if not opt.get("single_precision", True):
            raise NotImplementedError("the GPU solver computes in complex64 (single_precision=true)")
        self.opt = opt
        self.lam = float(opt.get("lambda", 1.0))
        self.callback_interval = opt.get("callback_interval", 50)
        self.energy_threshold = float(opt.get("energy_threshold", 1e-2))
        self.energy_calculation_interval = int(min(opt.get("energy_calculation_interval", 8),
                                                   self.callback_interval))
        self.sample = med = Medium(refractive_index, opt)
        self.grid = grid = med.grid
        self.N = list(grid.N) + [1]
        lo = np.array([b + 1 for b in med.Bl])
        hi = np.array([grid.N[d] - med.Br[d] for d in range(3)])
        if opt.get("roi") is None:                                       # Simulation.m:118-124
            self.roi = np.array([_make4(lo), _make4(hi)])
        else:
            r = np.asarray(opt["roi"])
            self.roi = np.array([_make4(r[0]), _make4(r[1])])
            self.roi[:, :3] += lo - 1
        dims = 2 if min(grid.N) == 1 else 3                              # PSTD.m:45-49
        self.dtmax = 2 / math.sqrt(dims) / math.pi * grid.dx * math.sqrt(med.e_r_min)
        self.dt = float(opt.get("dt_relative", 0.95)) * self.dtmax       # PSTD.m:51-52
        self.iterations_per_cycle = self.lam / self.dt                   # PSTD.m:53
        self.omega = 2 * math.pi / self.lam                              # PSTD.m:56
        e_r = med.e_r
        c2dt = self.dt ** 2 / e_r.real                                   # PSTD.m:57-62
        sdt = self.dt * self.omega * e_r.imag / e_r.real
        self.c1 = (sdt - 2) / (sdt + 2)
        self.c2 = 4 / (sdt + 2)
        self.c3 = 2 * c2dt / (sdt + 2)
        self.max_cycles = float(opt.get("max_cycles", math.inf)) + 100  # PSTD.m:69
        self._h = None

    # -- device handle ------------------------------------------------------------
    def _handle(self):
        if self._h is None:
            L = capi.lib()
            d = capi.pstd_desc_t()
            d.N[:] = [int(x) for x in self.grid.N]
            for c in range(4):
                d.roi[0][c] = int(self.roi[0, c])
                d.roi[1][c] = int(self.roi[1, c])
            d.iterations_per_cycle = float(self.iterations_per_cycle)
            d.energy_interval = int(self.energy_calculation_interval)
            cs = [np.asfortranarray(c.astype(np.float32)) for c in (self.c1, self.c2, self.c3)]
            ps = [np.ascontiguousarray(np.asarray(r, np.float32).ravel())
                  for r in (self.grid.py_range, self.grid.px_range, self.grid.pz_range)]
            h = C.c_void_p()
            capi.check(L.wsim_pstd_create(C.byref(h), C.byref(d), *[capi.ptr(a) for a in cs],
                                          *[capi.ptr(a) for a in ps]))
            self._h = h
        return self._h

    @property
    def fixed_launches(self) -> int:
        """Launches of one step that run plan-specialised pass kernels."""
        n = capi.lib().wsim_pstd_fixed_launches(self._handle())
        capi.check(n if n < 0 else 0)
        return n

    def close(self):
        if self._h is not None:
            capi.lib().wsim_pstd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def max_iterations(self) -> float:
        return math.ceil(self.max_cycles * self.iterations_per_cycle) if math.isfinite(self.max_cycles) \
            else math.inf

    def crop_field(self, E):
        r = self.roi
        return E[r[0, 0] - 1:r[1, 0], r[0, 1] - 1:r[1, 1], r[0, 2] - 1:r[1, 2], r[0, 3] - 1:r[1, 3]]

    def exec(self, source: Source, max_iterations: Optional[float] = None):
        """Simulation.exec (Simulation.m:138-190) with PSTD.run_algorithm:
        returns (E on the ROI times exp(-i angle A_next), state)."""
        t0 = time.perf_counter()
        src = source.shift(self.roi[0]).crop([1, 1, 1, 1], self.N)        # Simulation.m:154-156
        src.values = [v.astype(np.complex64) for v in src.values]
        L = capi.lib()
        h = self._handle()
        P = capi.Patches(src.positions, src.values)
        capi.check(L.wsim_pstd_set_source(h, P.n, P.arr))
        mi = self.max_iterations() if max_iterations is None else max_iterations
        n = C.c_int64()
        cv = C.c_int()
        t1 = time.perf_counter()
        capi.check(L.wsim_pstd_run(h, float(mi), self.energy_threshold, C.byref(n), C.byref(cv)))
        run_s = time.perf_counter() - t1
        E = np.empty(self.N, np.complex64, order="F")
        de = np.zeros(max(int(n.value), 1))
        capi.check(L.wsim_pstd_result(h, capi.ptr(E), de.ctypes.data_as(C.POINTER(C.c_double)), int(n.value)))
        K = int(n.value)
        st = PSTDState(it=K + 1, iterations=K, converged=cv.value == 1, diff_energy=de[:K], E_full=E,
                       run_time=run_s)
        A_next = default_source((K + 1) / self.iterations_per_cycle)   # PSTD.m:125-126
        ph = np.complex64(np.exp(-1j * np.angle(A_next)))
        Eout = np.ascontiguousarray(self.crop_field(E) * ph)
        st.time = time.perf_counter() - t0
        return Eout, st
