"""Host-side mirror of the reference's Simulation / Medium / SimulationGrid /
WaveSim / WaveSimVector / Source classes, running the loop on the GPU.

The constructor arithmetic follows the reference so that epsilon, k0c, gamma,
k02e and the wiggle descriptors -- which fix the iteration count -- are what
the MATLAB code computes; ``exec`` hands them to the HIP solver through the
C ABI instead of running ``run_algorithm``'s MATLAB loop.

Conventions are the reference's: arrays in [y, x, z, pol] index order,
positions and ROIs 1-based and inclusive.  Options use the reference's names
(``lambda``, ``pixel_size``, ``boundary_widths``, ``energy_threshold``,
``max_cycles``, ``callback_interval``, ``energy_calculation_interval``,
``ACC``/``wiggle``, ``epsilon``, ``roi``, ``boundary_type``).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

from . import capi, plans
from .solver import MBSSolver

ALPHA = 0.7  # f_mix.m:3


def _make4(v):
    v = [int(x) for x in np.ravel(v)]
    if len(v) > 4:
        raise ValueError("Position and size vectors can be 4-dimensional at most")
    return v + [1] * (4 - len(v))


def _prime_factors(n: int):
    out, p = [], 2
    while p * p <= n:
        while n % p == 0:
            out.append(p)
            n //= p
        p += 1
    if n > 1 or not out:
        out.append(n)
    return out


def efficient_size(s: int) -> int:
    """SimulationGrid.efficient_size (SimulationGrid.m:64-80): the next size
    whose largest prime factor is at most 11 and, with more than two factors,
    whose second largest is at most 5."""
    while True:
        f = _prime_factors(int(s))
        if not (f[-1] > 11 or (len(f) > 2 and f[-2] > 5)):
            return int(s)
        s += 1


def symrange(n: int) -> np.ndarray:
    """SimulationGrid.symrange (SimulationGrid.m:55-63)."""
    return np.fft.fftshift(np.arange(n) - n // 2).astype(np.float64)


class SimulationGrid:
    """SimulationGrid (SimulationGrid.m:20-52); entries are [y, x, z]."""

    def __init__(self, min_size, dx, periodic):
        min_size = [int(math.ceil(s)) for s in min_size]
        self.N = [s if periodic[d] else efficient_size(s) for d, s in enumerate(min_size)]
        self.periodic = list(periodic)
        self.padding = [self.N[d] - min_size[d] for d in range(3)]
        self.dx = float(dx)
        Ny, Nx, Nz = self.N
        self.y_range = np.arange(Ny) * dx
        self.x_range = np.arange(Nx) * dx
        self.z_range = np.arange(Nz) * dx
        self.dpy = 2 * math.pi / (dx * Ny)
        self.dpx = 2 * math.pi / (dx * Nx)
        self.dpz = 2 * math.pi / (dx * Nz)
        self.py_range = self.dpy * symrange(Ny)
        self.px_range = self.dpx * symrange(Nx)
        self.pz_range = self.dpz * symrange(Nz)


class Medium:
    """Medium (Medium.m:15-72) with replicate padding (Medium.m:126-132).
    Only the 'ARL' boundary is supported (the 'PBL' types are deprecated in
    the reference, Medium.m:86)."""

    def __init__(self, refractive_index, opt: dict):
        n = np.asarray(refractive_index)
        if n.ndim == 1:
            n = n[:, None]
        if n.ndim > 3:
            raise ValueError("refractive index must be 1-D, 2-D or 3-D")
        if np.min(np.imag(n)) < 0:
            raise ValueError("Medium cannot have gain, imaginary part of refractive index should be positive")
        self.boundary_type = opt.get("boundary_type", "ARL")
        if self.boundary_type != "ARL":
            raise NotImplementedError("only the ARL boundary type is supported")
        bw = [int(b) for b in np.ravel(opt["boundary_widths"])]
        if not (n.ndim <= len(bw) <= 3):
            raise ValueError("boundary_widths must have one entry per medium dimension (at most 3)")
        e_r = n.astype(np.complex128) ** 2
        self.e_r_min = float(e_r.real.min())
        self.e_r_max = float(e_r.real.max())
        self.e_r_center = (self.e_r_min + self.e_r_max) / 2
        sz = list(n.shape) + [1] * (3 - n.ndim)
        bw2 = [2 * b for b in bw] + [0] * (3 - len(bw))
        self.grid = SimulationGrid([sz[d] + bw2[d] for d in range(3)], opt["pixel_size"],
                                   [b == 0 for b in bw2])
        N = self.grid.N
        self.Bl = [(N[d] - sz[d] + 1) // 2 for d in range(3)]
        self.Br = [(N[d] - sz[d]) // 2 for d in range(3)]
        self.e_r = np.pad(e_r.reshape(sz), [(self.Bl[d], self.Br[d]) for d in range(3)], mode="edge")
        self.boundary_widths = bw + [0] * (3 - len(bw))


def _nuttall(L: int) -> np.ndarray:
    """4-term symmetric Nuttall window (coefficients of MATLAB nuttallwin)."""
    if L == 1:
        return np.ones(1)
    t = 2 * np.pi * np.arange(L) / (L - 1)
    return 0.3635819 - 0.4891775 * np.cos(t) + 0.1365995 * np.cos(2 * t) - 0.0106411 * np.cos(3 * t)


def _edge_filter(N, bl, br, L, three_d):
    """One ARL edge filter (WaveSimBase.m:230-252)."""
    roi = N - bl - br
    if three_d:
        win = lambda B: (np.arange(1, B + 1) - 0.21) / (B + 0.66)
    else:
        win = lambda B: _nuttall(2 * L - 1)[:B]
    smooth = np.concatenate([[0.0], win(L - 1)]) if br == bl else win(L)
    return np.concatenate([np.zeros(bl - L), smooth, np.ones(roi), smooth[::-1], np.zeros(br - L)])


def _wiggle_set(flags):
    """private/wiggle_perm.m: +-1 sign table, columns reordered to reduce the
    second-order wrap-around artefact."""
    n = int(sum(flags))
    nw = 2 ** n
    cols = np.arange(1, nw + 1)
    table = np.zeros((3, nw))
    rows = iter([(-1.0) ** (np.ceil(cols / 2.0 ** r) + 1) for r in range(n)])
    for d in range(3):
        if flags[d]:
            table[d] = next(rows)
    order = {2: [1, 2, 4, 3], 3: [1, 2, 3, 5, 4, 6, 7, 8]}.get(n, list(range(1, nw + 1)))
    return table[:, [i - 1 for i in order]]


@dataclass
class SimState:
    it: int = 0
    iterations: int = 0
    converged: bool = False
    diff_energy: np.ndarray = field(default_factory=lambda: np.zeros(0))
    energy_every: np.ndarray = field(default_factory=lambda: np.zeros(0))
    time: float = 0.0
    E_primed: Optional[np.ndarray] = None
    source_energy: float = 0.0


class Source:
    """Source (Source.m:1-174): a list of (1-based [y,x,z,pol] position, value
    array) patches.  1-D values are column vectors (n x 1), as ``ones(N,1)``."""

    def __init__(self, value=None, position=(1, 1, 1, 1)):
        self.positions, self.values = [], []
        if value is not None:
            self.positions.append(_make4(position))
            self.values.append(self._as4(value))

    @staticmethod
    def _as4(v):
        v = np.asarray(v)
        if v.ndim == 0:
            return v.reshape(1, 1, 1, 1)
        if v.ndim == 1:
            return v.reshape(-1, 1, 1, 1)
        return v.reshape(list(v.shape) + [1] * (4 - v.ndim))

    def __add__(self, other: "Source") -> "Source":
        s = Source()
        s.positions = self.positions + other.positions
        s.values = self.values + other.values
        return s

    def energy(self) -> float:
        return float(sum(np.sum(np.abs(v.astype(np.complex128)) ** 2) for v in self.values))

    def shift(self, pos) -> "Source":
        s = Source()
        s.positions = [list(np.asarray(p) + np.asarray(pos) - 1) for p in self.positions]
        s.values = list(self.values)
        return s

    def crop(self, lo, hi) -> "Source":
        """Keep the parts of the patches inside [lo, hi] (Source.m:81-103)."""
        s = Source()
        lo, hi = np.asarray(lo), np.asarray(hi)
        for p, v in zip(self.positions, self.values):
            p = np.asarray(p)
            tl = np.maximum(lo, p)
            br = np.minimum(hi, p + np.asarray(v.shape) - 1)
            if np.any(tl > br):
                continue
            a, b = tl - p, br - p + 1
            s.positions.append(list(tl))
            s.values.append(v[a[0]:b[0], a[1]:b[1], a[2]:b[2], a[3]:b[3]])
        return s


class WaveSimBase:
    """WaveSimBase construction (WaveSimBase.m:40-82, Simulation.m:80-136)
    and ``exec`` (Simulation.m:138-190) with the loop on the GPU."""

    vector = False

    def __init__(self, refractive_index, options: dict):
        opt = dict(options)
        if "wiggle" in opt:
            opt["ACC"] = opt["wiggle"]
        # single_precision (Simulation.m:31): complex64 fields (default) or
        # complex128 (data_array, Simulation.m:221-252)
        self.single = bool(opt.get("single_precision", True))
        self.cdtype = np.complex64 if self.single else np.complex128
        self.rdtype = np.float32 if self.single else np.float64
        self.opt = opt
        self.lam = float(opt.get("lambda", 1.0))
        self.callback_interval = opt.get("callback_interval", 50)
        self.energy_threshold = float(opt.get("energy_threshold", 1e-2))
        self.max_cycles = float(opt.get("max_cycles", math.inf))
        self.energy_calculation_interval = int(min(opt.get("energy_calculation_interval", 8),
                                                   self.callback_interval))
        self.callback = opt.get("callback")
        self.sample = med = Medium(refractive_index, opt)
        self.grid = grid = med.grid
        self.N = list(grid.N) + [1]
        lo = np.array([b + 1 for b in med.Bl])
        hi = np.array([grid.N[d] - med.Br[d] for d in range(3)])
        if opt.get("roi") is None:
            self.roi = np.array([_make4(lo), _make4(hi)])
        else:
            r = np.asarray(opt["roi"])
            self.roi = np.array([_make4(r[0]), _make4(r[1])])
            self.roi[:, :3] += lo - 1
        k0 = 2 * math.pi / self.lam
        k0c = math.sqrt(med.e_r_center) * k0
        if opt.get("epsilon") is not None:
            eps = float(opt["epsilon"])
        else:
            eps = max(float(np.max(np.abs(med.e_r * k0 ** 2 - k0c ** 2))) / 0.95,
                      float(opt.get("epsilonmin", 3)))
        self.k0, self.k0c = k0, k0c
        self.iterations_per_cycle = self.lam / (2 * k0c / eps)
        V = med.e_r * k0 ** 2 - k0c ** 2 - 1.0j * eps
        self.filters = [None, None, None]
        if not all(b == 0 for b in med.Bl):
            three_d = all(n > 1 for n in grid.N)
            for d in range(3):
                if med.Bl[d] > 0:
                    f = _edge_filter(grid.N[d], med.Bl[d], med.Br[d], med.boundary_widths[d], three_d)
                    self.filters[d] = f
                    shape = [1, 1, 1]
                    shape[d] = -1
                    V = V * f.reshape(shape)
        self.gamma = (1.0j / eps * V).astype(self.cdtype)                 # WaveSimBase.m:73
        self.k02e = self.cdtype(k0c ** 2 / eps + 1.0j)                    # WaveSimBase.m:77
        self.epsilon = self.rdtype(eps)                                   # WaveSimBase.m:78
        acc = np.ravel(opt.get("ACC", True))
        if acc.size == 1:
            acc = [not p for p in grid.periodic] if bool(acc[0]) else [False] * 3
        self.ACC = [bool(a) for a in acc]
        self.wiggles = self._wiggles(_wiggle_set(self.ACC))
        if self.vector:
            if opt.get("roi") is None:
                self.roi[1, 3] = 3
            self.N[3] = 3
        self._solver = None
        self._solver_nrhs = 0

    def _wiggles(self, table):
        g = self.grid
        sq = np.sqrt(self.epsilon)  # epsilon in the field's precision (WaveSimBase.m:78, 307-309)
        rdt, cdt = self.rdtype, self.cdtype
        out = []
        for w in range(table.shape[1]):
            wy, wx, wz = table[:, w]
            def coord(pr, dp, s):
                return ((pr - dp * s / 4).astype(rdt) / sq).astype(rdt)
            def ramp(rng, s):
                return np.exp(1.0j * math.pi / 2 * s * (rng / g.dx / len(rng))).astype(cdt)
            out.append(dict(pxe=coord(g.px_range, g.dpx, wx), pye=coord(g.py_range, g.dpy, wy),
                            pze=coord(g.pz_range, g.dpz, wz), gx=ramp(g.x_range, wx),
                            gy=ramp(g.y_range, wy), gz=ramp(g.z_range, wz)))
        return out

    # -- solver ---------------------------------------------------------------
    def solver(self, nrhs: int = 1) -> MBSSolver:
        if self._solver is None or self._solver_nrhs != nrhs:
            if self._solver is not None:
                self._solver.close()
            self._solver = MBSSolver(self.N, self.gamma, self.wiggles, float(self.epsilon),
                                     complex(self.k02e), self.roi, self.energy_calculation_interval,
                                     nrhs=nrhs, dtype=self.cdtype)
            self._solver_nrhs = nrhs
            plan = plans.lookup(self.N, nrhs, len(self.wiggles), f64=not self.single)
            if plan:
                try:
                    self._solver.plan = plan
                except capi.WsimError as e:
                    # a plan that needs kernels this library does not run
                    # (plan-specialised kernels turned off): the library keeps
                    # the plan it had, which computes the same field
                    if e.code != capi.WSIM_ERR_UNSUPPORTED:
                        raise
        return self._solver

    def adopt_solver(self, other: "WaveSimBase"):
        """Take over another simulation's solver (same grid, wiggle count and
        ROI), loading this medium into it instead of allocating a new one."""
        sv = other._solver
        if not self.single or (sv is not None and sv.f64):
            raise NotImplementedError("complex128 solvers take no new medium: create a solver per medium")
        if sv is None or list(sv.N) != list(self.N) or sv.nw != len(self.wiggles) \
                or not np.array_equal(other.roi, self.roi):
            raise ValueError("solver shape does not match this simulation")
        sv.set_medium(self.gamma, self.wiggles, float(self.epsilon), complex(self.k02e))
        other._solver = None
        self._solver = sv
        self._solver_nrhs = other._solver_nrhs
        return sv

    def tune(self, nrhs: int = 1, reps: int = 4):
        """Autotune the FFT plan of this problem (radix order and tile width
        per axis) and remember it for later solvers of the same shape."""
        sv = self.solver(nrhs)
        plan, ms = sv.tune(reps)
        plans.store(self.N, nrhs, len(self.wiggles), plan, ms, f64=not self.single)
        return plan, ms

    def _prepare(self, source: Source) -> Source:
        """Simulation.m:154-156: shift to the ROI origin and crop to the grid."""
        src = source.shift(self.roi[0]).crop([1, 1, 1, 1], self.N)
        src.values = [v.astype(self.cdtype) for v in src.values]
        return src

    def max_iterations(self) -> int:
        if math.isfinite(self.max_cycles):
            return int(math.ceil(self.max_cycles * self.iterations_per_cycle))
        return 100000  # WaveSimBase.m:127-131

    def crop_field(self, E):
        r = self.roi
        return E[r[0, 0] - 1:r[1, 0], r[0, 1] - 1:r[1, 1], r[0, 2] - 1:r[1, 2], r[0, 3] - 1:r[1, 3]]

    def exec(self, source: Source, max_iterations: Optional[int] = None):
        """Solve for one source: returns (E cropped to the ROI, state)."""
        E, states = self.exec_batch([source], max_iterations)
        return E[0], states[0]

    def exec_batch(self, sources: List[Source], max_iterations: Optional[int] = None):
        """Solve several sources on the same medium in one batched launch; each
        result equals exec() of that source alone."""
        t0 = time.perf_counter()
        prepared = [self._prepare(s) for s in sources]
        sv = self.solver(len(sources))
        for r, src in enumerate(prepared):
            sv.set_source(r, src.positions, src.values)
        mi = int(max_iterations) if max_iterations is not None else self.max_iterations()
        cb = None
        if self.callback is not None:
            cb = lambda it, hist: self.callback(it, hist)
        sv.solve(mi, self.energy_threshold, callback=cb,
                 callback_interval=float(self.callback_interval))
        return self._finish(sv, prepared, t0)

    def propagate(self, E, wiggle: int = 0):
        return self.solver().propagate(wiggle, E)

    def _finish(self, sv, prepared, t0):
        Es, states = [], []
        g4 = self.gamma[..., None]
        for r, src in enumerate(prepared):
            res = sv.result(r)
            Ep = sv.field(r)
            with np.errstate(divide="ignore", invalid="ignore"):
                E = self.crop_field(Ep / g4)
            st = SimState(it=res.iterations + 1, iterations=res.iterations, converged=res.converged,
                          diff_energy=res.diff_energy, energy_every=res.energy_every,
                          time=time.perf_counter() - t0, E_primed=Ep,
                          source_energy=src.energy())
            Es.append(np.ascontiguousarray(E))
            states.append(st)
        return Es, states


def exec_many(problems, max_iterations: Optional[int] = None):
    """Solve several independent problems -- (simulation, source) pairs with
    any grids and media -- in one schedule (a group of solvers sharing
    launches).  Returns [(E, state), ...]; each equals sim.exec(source).
    The reference solves such sets as a loop of exec calls
    (examples/pulse2d.m:31-40, examples/phase_conjugation.m:48-56)."""
    from .solver import SolverGroup
    t0 = time.perf_counter()
    sims = [p[0] for p in problems]
    if len(set(id(s) for s in sims)) != len(sims):
        raise ValueError("exec_many: each simulation may appear once (use exec_batch for several "
                         "sources on one medium)")
    prepared, svs = [], []
    for sim, src in problems:
        ps = sim._prepare(src)
        sv = sim.solver(1)
        sv.set_source(0, ps.positions, ps.values)
        prepared.append(ps)
        svs.append(sv)
    mi = [int(max_iterations) if max_iterations is not None else s.max_iterations() for s in sims]
    with SolverGroup(svs) as g:
        g.solve(mi, [s.energy_threshold for s in sims])
    out = []
    for sim, sv, ps in zip(sims, svs, prepared):
        Es, states = sim._finish(sv, [ps], t0)
        out.append((Es[0], states[0]))
    return out


class WaveSim(WaveSimBase):
    """Scalar Helmholtz MBS (WaveSim.m)."""
    vector = False


class WaveSimVector(WaveSimBase):
    """Vector Maxwell MBS (WaveSimVector.m): N(4) = 3, component 1 = x."""
    vector = True
