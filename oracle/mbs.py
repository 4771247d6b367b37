"""numpy restatement of the reference MBS solver (setup + MATLAB-branch loop).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.

Array conventions follow the reference: index order ``[y, x, z, pol]``
(``Simulation.m:22-26``, ``WaveSimVector.m:17``), positions and ROIs 1-based and
inclusive.  Arrays here are numpy arrays of shape ``(Ny, Nx, Nz, Npol)`` in
MATLAB index order; ``np.asfortranarray`` of them is the MATLAB memory layout
(y contiguous) that the C ABI takes.

Precision: ``single=True`` (the reference default, ``Simulation.m:31``)
reproduces MATLAB's single/double mixing rules where the reference mixes them
(``data_array`` casts, ``Simulation.m:221-252``; wiggle coordinates divided by
the already-single epsilon, ``WaveSimBase.m:78-81, 307-309``).  ``single=False``
runs the same algorithm in complex128.

Pins (see tests/test_oracle_*.py):
  factor / efficient_size / symrange  -- hand tables (closed form).
  nuttallwin                          -- scipy.signal.windows.nuttall and the
                                          published 4-term coefficients.
  Medium padding                      -- numpy.pad(mode='edge') + hand Bl/Br.
  epsilon / V / gamma / k02e          -- closed forms for homogeneous media.
  apply_edge_filters                  -- hand-built small filters.
  wiggle_perm / wiggle_descriptor     -- hand tables / closed form.
  propagate (scalar)                  -- dense DFT-matrix brute force on tiny
                                          grids; periodic homogeneous closed form.
  propagate (vector)                  -- invariant p.G(f) = -(p.f)/k02e and
                                          reduction to the scalar case for
                                          transverse fields.
  run loop + stopping rule            -- periodic closed-form solution,
                                          Helmholtz fixed point, linearity,
                                          hand-traced stop iterations.
  whole solver with ARL + ACC         -- 1-D analytic solution
                                          (homogeneous_medium_analytic_solution)
                                          and the 2-D Hankel solution.
Parity unpinned (needs MATLAB itself): MATLAB's float32 ``sum`` order in
``Simulation.energy`` (this oracle accumulates in float64) and FFTW/cuFFT
last-bit rounding.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np
import scipy.fft as sfft

ALPHA = 0.7  # private/f_mix.m:3 (hard-wired in the reference)


# ----------------------------------------------------------------------------
# small helpers (Simulation.m:352-379)
# ----------------------------------------------------------------------------
def make3(sz, pad):
    """Simulation.make3 (Simulation.m:352-361): pad a size vector to 3 entries."""
    sz = [int(s) if float(s).is_integer() else s for s in np.ravel(sz)]
    while len(sz) < 3:
        sz.append(pad)
    return sz


def make4(sz):
    """Simulation.make4 (Simulation.m:363-379): pad a size/position vector to 4."""
    sz = list(np.ravel(sz))
    if len(sz) > 4:
        raise ValueError("Position and size vectors can be 4-dimensional at most")
    while len(sz) < 4:
        sz.append(1)
    return [int(s) for s in sz]


def matlab_size4(a) -> List[int]:
    """size(a) made 4-long, with numpy arrays read in MATLAB index order.

    A 0-d value is a 1x1 MATLAB scalar; a 1-d numpy array is taken as a MATLAB
    column vector (n x 1), as ``ones(N,1)`` is in the reference examples.
    """
    a = np.asarray(a)
    if a.ndim == 0:
        return [1, 1, 1, 1]
    if a.ndim == 1:
        return [a.shape[0], 1, 1, 1]
    return make4(a.shape)


def as4d(a) -> np.ndarray:
    """Reshape a value array to its 4-D MATLAB shape ``matlab_size4(a)``."""
    a = np.asarray(a)
    return a.reshape(matlab_size4(a), order="C") if a.ndim <= 1 else a.reshape(make4(a.shape))


# ----------------------------------------------------------------------------
# SimulationGrid (SimulationGrid.m:20-80)
# ----------------------------------------------------------------------------
def factor(n: int) -> List[int]:
    """MATLAB ``factor``: prime factors in ascending order, ``factor(1) == [1]``."""
    n = int(n)
    if n == 1:
        return [1]
    out, p = [], 2
    while p * p <= n:
        while n % p == 0:
            out.append(p)
            n //= p
        p += 1
    if n > 1:
        out.append(n)
    return out


def efficient_size(s: int) -> int:
    """SimulationGrid.efficient_size (SimulationGrid.m:64-80), one entry."""
    s = int(s)
    f = factor(s)
    while f[-1] > 11 or (len(f) > 2 and f[-2] > 5):
        s += 1
        f = factor(s)
    return s


def symrange(n: int) -> np.ndarray:
    """SimulationGrid.symrange (SimulationGrid.m:55-63): fftshift((0:N-1)-floor(N/2))."""
    return np.fft.fftshift(np.arange(n) - n // 2).astype(np.float64)


@dataclass
class Grid:
    """SimulationGrid (SimulationGrid.m:20-52); entries are [y, x, z]."""
    N: List[int]
    dx: float
    periodic: List[bool]
    padding: List[int]
    x_range: np.ndarray
    y_range: np.ndarray
    z_range: np.ndarray
    dpx: float
    dpy: float
    dpz: float
    px_range: np.ndarray
    py_range: np.ndarray
    pz_range: np.ndarray


def simulation_grid(min_size, dx, periodic) -> Grid:
    """SimulationGrid constructor (SimulationGrid.m:20-52)."""
    min_size = [int(math.ceil(s)) for s in min_size]
    N = list(min_size)
    for d in range(3):
        if not periodic[d]:
            N[d] = efficient_size(min_size[d])
    padding = [N[d] - min_size[d] for d in range(3)]
    x_range = np.arange(N[1]) * dx
    y_range = np.arange(N[0]) * dx
    z_range = np.arange(N[2]) * dx
    dpx = 2 * math.pi / (dx * N[1])
    dpy = 2 * math.pi / (dx * N[0])
    dpz = 2 * math.pi / (dx * N[2])
    return Grid(N=N, dx=dx, periodic=list(periodic), padding=padding,
                x_range=x_range, y_range=y_range, z_range=z_range,
                dpx=dpx, dpy=dpy, dpz=dpz,
                px_range=dpx * symrange(N[1]), py_range=dpy * symrange(N[0]),
                pz_range=dpz * symrange(N[2]))


# ----------------------------------------------------------------------------
# Medium (Medium.m:15-72, 126-132)
# ----------------------------------------------------------------------------
@dataclass
class MediumData:
    e_r: np.ndarray        # padded permittivity, (Ny, Nx, Nz) complex128
    e_r_min: float
    e_r_max: float
    e_r_center: float
    Bl: List[int]
    Br: List[int]
    boundary_type: str
    grid: Grid


def medium(refractive_index, opt: dict) -> MediumData:
    """Medium constructor (Medium.m:15-72) with the 'replicate' padding of
    Medium.extrapolate (Medium.m:126-132).  The deprecated PBL boundary
    (Medium.m:83-123) is not restated."""
    n = np.asarray(refractive_index)
    if n.ndim == 1:
        n = n.reshape(-1, 1)
    if np.min(np.imag(n)) < 0:
        raise ValueError("Medium cannot have gain, imaginary part of refractive index should be positive")
    btype = opt.get("boundary_type", "ARL")
    if btype != "ARL":
        raise NotImplementedError("only the ARL boundary is restated (PBL is deprecated in the reference)")
    bwidths = list(np.ravel(opt["boundary_widths"]))
    assert len(bwidths) >= n.ndim and len(bwidths) <= 3
    e_r = n.astype(np.complex128) ** 2
    e_r_min = float(np.min(e_r.real))
    e_r_max = float(np.max(e_r.real))
    e_r_center = (e_r_min + e_r_max) / 2
    sz = make3(n.shape, 1)
    bw = make3([2 * b for b in bwidths], 0)
    grid = simulation_grid([sz[d] + bw[d] for d in range(3)], opt["pixel_size"],
                           [bw[d] == 0 for d in range(3)])
    Bl = [int(math.ceil((grid.N[d] - sz[d]) / 2)) for d in range(3)]
    Br = [int(math.floor((grid.N[d] - sz[d]) / 2)) for d in range(3)]
    e3 = e_r.reshape(sz)
    e3 = np.pad(e3, [(Bl[d], Br[d]) for d in range(3)], mode="edge")
    return MediumData(e_r=e3, e_r_min=e_r_min, e_r_max=e_r_max, e_r_center=e_r_center,
                      Bl=Bl, Br=Br, boundary_type=btype, grid=grid)


# ----------------------------------------------------------------------------
# windows and wiggles
# ----------------------------------------------------------------------------
NUTTALL = (0.3635819, 0.4891775, 0.1365995, 0.0106411)


def nuttallwin(L: int) -> np.ndarray:
    """MATLAB ``nuttallwin(L)``: symmetric 4-term Nuttall window (published coefficients)."""
    if L == 1:
        return np.ones(1)
    n = np.arange(L)
    a0, a1, a2, a3 = NUTTALL
    t = 2 * np.pi * n / (L - 1)
    return a0 - a1 * np.cos(t) + a2 * np.cos(2 * t) - a3 * np.cos(3 * t)


def edge_filters(med: MediumData, opt: dict):
    """The per-dimension ARL filters of apply_edge_filters (WaveSimBase.m:217-258).

    Returns a list of 3 entries, each a 1-D filter of length N[d] or None."""
    filters = [None, None, None]
    if med.boundary_type != "ARL" or all(b == 0 for b in med.Bl):
        return filters
    N = med.grid.N
    bwidths = list(np.ravel(opt["boundary_widths"]))
    for d in range(3):
        bl, br = med.Bl[d], med.Br[d]
        roi_size = N[d] - bl - br
        if bl > 0:
            L = int(bwidths[d])
            if all(x > 1 for x in N):           # 3-D: linear window
                window = lambda B: (np.arange(1, B + 1) - 0.21) / (B + 0.66)
            else:                               # 1-D / 2-D: Nuttall window
                window = lambda B, L=L: nuttallwin(2 * L - 1)[:B]
            if br == bl:
                smooth = np.concatenate([[0.0], window(L - 1)])
            else:
                smooth = window(L)
            filt = np.concatenate([np.zeros(bl - L), smooth, np.ones(roi_size),
                                   smooth[::-1], np.zeros(br - L)])
            filters[d] = filt
    return filters


def wiggle_perm(flags: Sequence[bool]) -> np.ndarray:
    """private/wiggle_perm.m:1-28 -- sign table (3 x 2^n), columns reordered."""
    flags = [bool(f) for f in flags]
    n = sum(flags)
    nw = 2 ** n
    ws = np.zeros((len(flags), nw))
    cols = np.arange(1, nw + 1)
    rows = [(-1.0) ** (np.ceil(cols / 2.0 ** r) + 1) for r in range(n)]
    ri = 0
    for d, f in enumerate(flags):
        if f:
            ws[d] = rows[ri]
            ri += 1
    ind = list(range(1, nw + 1))
    if n == 2:
        ind = [1, 2, 4, 3]
    elif n == 3:
        ind = [1, 2, 3, 5, 4, 6, 7, 8]
    return ws[:, [i - 1 for i in ind]]


# ----------------------------------------------------------------------------
# the simulation object (WaveSimBase.m:40-82, Simulation.m:80-136)
# ----------------------------------------------------------------------------
@dataclass
class Sim:
    N: List[int]              # [Ny, Nx, Nz, Npol]
    roi: np.ndarray           # 2x4 int, 1-based inclusive
    gamma: np.ndarray         # (Ny, Nx, Nz) complex
    epsilon: float            # single (or double) epsilon as stored
    k02e: complex
    wiggles: list             # list of dicts: pxe, pye, pze (real), gx, gy, gz (complex)
    wiggle_set: np.ndarray
    ACC: List[bool]
    k0: float
    k0c: float
    iterations_per_cycle: float
    energy_threshold: float
    energy_calculation_interval: int
    callback_interval: float
    max_cycles: float
    single: bool
    vector: bool
    medium: MediumData
    V: np.ndarray             # filtered potential (double), for physics checks
    filters: list

    @property
    def cdtype(self):
        return np.complex64 if self.single else np.complex128

    @property
    def rdtype(self):
        return np.float32 if self.single else np.float64


def make_sim(refractive_index, opt: dict, vector: bool = False) -> Sim:
    """WaveSim / WaveSimVector construction (WaveSimBase.m:40-82,
    Simulation.m:80-136, WaveSimVector.m:12-18)."""
    opt = dict(opt)
    if "wiggle" in opt:                                   # WaveSimBase.m:48-50
        opt["ACC"] = opt["wiggle"]
    single = bool(opt.get("single_precision", True))
    rdt = np.float32 if single else np.float64
    cdt = np.complex64 if single else np.complex128
    lam = float(opt.get("lambda", 1.0))
    callback_interval = opt.get("callback_interval", 50)
    eci = opt.get("energy_calculation_interval", 8)
    eci = min(eci, callback_interval)                     # Simulation.m:135
    med = medium(refractive_index, opt)
    grid = med.grid
    N = list(grid.N) + [1]
    default_roi = np.array([[b + 1 for b in med.Bl], [grid.N[d] - med.Br[d] for d in range(3)]])
    if opt.get("roi") is None:
        roi = np.array([make4(default_roi[0]), make4(default_roi[1])])
    else:
        r = np.asarray(opt["roi"])
        roi = np.array([make4(r[0]), make4(r[1])])
        roi[:, :3] = roi[:, :3] + default_roi[0] - 1
    # WaveSimBase constructor
    k0 = 2 * math.pi / lam
    k0c = math.sqrt(med.e_r_center) * k0
    if "epsilon" in opt and opt["epsilon"] is not None:
        epsilon = float(opt["epsilon"])
    else:                                                 # WaveSimBase.m:201-215
        vabs = float(np.max(np.abs(med.e_r * k0 ** 2 - k0c ** 2))) / 0.95
        epsilon = max(vabs, float(opt.get("epsilonmin", 3)))
    ipc = lam / (2 * k0c / epsilon)
    V = med.e_r * k0 ** 2 - k0c ** 2 - 1.0j * epsilon     # WaveSimBase.m:69
    filters = edge_filters(med, opt)
    for d in range(3):
        if filters[d] is not None:
            shape = [1, 1, 1]
            shape[d] = -1
            V = V * filters[d].reshape(shape)
    gamma = (1.0j / epsilon * V).astype(cdt)               # WaveSimBase.m:73
    k02e = cdt(k0c ** 2 / epsilon + 1.0j)                  # WaveSimBase.m:77
    eps_s = rdt(epsilon)                                   # WaveSimBase.m:78
    # compute_wiggles (WaveSimBase.m:272-298)
    acc = opt.get("ACC", True)
    acc = np.ravel(acc)
    if acc.size == 1:
        acc = [not p for p in grid.periodic] if bool(acc[0]) else [False] * 3
    else:
        acc = [bool(a) for a in acc]
    ws = wiggle_perm(acc)
    wiggles = []
    sq = np.sqrt(eps_s)                                    # sqrt of the (single) epsilon
    for w in range(ws.shape[1]):
        wig = ws[:, w]                                     # [y; x; z]
        def coord(pr, dp, wv):
            num = (pr - dp * wv / 4).astype(rdt)
            return (num / sq).astype(rdt)
        def ramp(rng, wv):
            ph = 1.0j * math.pi / 2 * wv * (rng / grid.dx / len(rng))
            return np.exp(ph).astype(cdt)
        wiggles.append(dict(
            pxe=coord(grid.px_range, grid.dpx, wig[1]),
            pye=coord(grid.py_range, grid.dpy, wig[0]),
            pze=coord(grid.pz_range, grid.dpz, wig[2]),
            gx=ramp(grid.x_range, wig[1]),
            gy=ramp(grid.y_range, wig[0]),
            gz=ramp(grid.z_range, wig[2])))
    if vector:                                             # WaveSimVector.m:14-17
        if opt.get("roi") is None:
            roi[1, 3] = 3
        N[3] = 3
    return Sim(N=N, roi=roi, gamma=gamma, epsilon=eps_s, k02e=k02e, wiggles=wiggles,
               wiggle_set=ws, ACC=acc, k0=k0, k0c=k0c, iterations_per_cycle=ipc,
               energy_threshold=float(opt.get("energy_threshold", 1e-2)),
               energy_calculation_interval=int(eci), callback_interval=callback_interval,
               max_cycles=float(opt.get("max_cycles", math.inf)), single=single,
               vector=vector, medium=med, V=V, filters=filters)


# ----------------------------------------------------------------------------
# Source (Source.m:24-172) -- a list of (position[4], value-4d) patches
# ----------------------------------------------------------------------------
@dataclass
class Source:
    positions: list = field(default_factory=list)
    values: list = field(default_factory=list)

    @staticmethod
    def point(value, position) -> "Source":
        """Source(value, position) (Source.m:24-46)."""
        return Source([make4(position)], [as4d(value)])

    def __add__(self, other: "Source") -> "Source":
        """Source.plus (Source.m:48-55)."""
        return Source(self.positions + other.positions, self.values + other.values)

    def energy(self) -> float:
        """Source.energy (Source.m:57-67)."""
        return float(sum(np.sum(np.abs(np.asarray(v, np.complex128)) ** 2) for v in self.values))

    def shift(self, pos) -> "Source":
        """Source.shift (Source.m:69-79)."""
        pos = np.asarray(pos)
        return Source([list(np.asarray(p) + pos - 1) for p in self.positions], list(self.values))

    def crop(self, roi) -> "Source":
        """Source.crop (Source.m:81-103)."""
        roi = np.asarray(roi)
        P, Vs = [], []
        for p, v in zip(self.positions, self.values):
            p = np.asarray(p)
            sz = np.asarray(v.shape)
            tlt = np.maximum(roi[0], p)
            brt = np.minimum(roi[1], p + sz - 1)
            if np.any(tlt > brt):
                continue
            tls = tlt - p
            brs = brt - p + 1
            Vs.append(v[tls[0]:brs[0], tls[1]:brs[1], tls[2]:brs[2], tls[3]:brs[3]])
            P.append(list(tlt))
        return Source(P, Vs)

    def cast(self, dtype) -> "Source":
        """data_array on a Source (Simulation.m:245-248 via Source.cast)."""
        out = []
        for v in self.values:
            v = np.asarray(v)
            if np.iscomplexobj(v):
                out.append(v.astype(np.complex64 if dtype == np.float32 else np.complex128))
            else:
                out.append(v.astype(dtype))
        return Source(list(self.positions), out)

    def add_to(self, E: np.ndarray, A) -> np.ndarray:
        """Source.add_to (Source.m:139-172): E(patch) += A*val for every patch."""
        for p, v in zip(self.positions, self.values):
            t = [int(x) - 1 for x in p]
            s = v.shape
            E[t[0]:t[0] + s[0], t[1]:t[1] + s[1], t[2]:t[2] + s[2], t[3]:t[3] + s[3]] += A * v
        return E

    def dense(self, N, dtype) -> np.ndarray:
        """All patches summed onto a zero grid of size N (test helper)."""
        out = np.zeros(N, dtype)
        return self.add_to(out, dtype(1))


def prepare_source(sim: Sim, source: Source) -> Source:
    """exec's source handling (Simulation.m:154-156): cast, shift to the ROI
    origin, crop to the grid."""
    return source.cast(sim.rdtype).shift(sim.roi[0]).crop(np.array([[1, 1, 1, 1], sim.N]))


# ----------------------------------------------------------------------------
# pointwise kernels (private/*.m) and propagate (WaveSim.m / WaveSimVector.m)
# ----------------------------------------------------------------------------
def f_wiggle(E, g1, g2, g3):
    """private/f_wiggle.m:5 -- E .* g1 .* g2 .* g3 (broadcast [y,x,z,pol])."""
    return E * g1 * g2 * g3


def f_mix(Ediff, Eprop, gamma, single=True):
    """private/f_mix.m:3-4 -- (1 - a*gamma) dE - 1i*a*gamma^2 Eprop."""
    if single:
        a = np.float32(ALPHA)
        ai = np.complex64(1j * ALPHA)
    else:
        a = ALPHA
        ai = 1j * ALPHA
    return (1 - a * gamma) * Ediff - (ai * (gamma * gamma)) * Eprop


def _fftn(E, inverse=False, workers=None):
    axes = (0, 1, 2)
    if inverse:
        return sfft.ifftn(E, axes=axes, workers=workers)
    return sfft.fftn(E, axes=axes, workers=workers)


def _ramps(wig, conj=False):
    gx = wig["gx"].reshape(1, -1, 1, 1)
    gy = wig["gy"].reshape(-1, 1, 1, 1)
    gz = wig["gz"].reshape(1, 1, -1, 1)
    if conj:
        return np.conj(gx), np.conj(gy), np.conj(gz)
    return gx, gy, gz


def propagate_scalar(sim: Sim, E, wig, workers=None):
    """WaveSim.propagate (WaveSim.m:15-36) with mfft/mifft (WaveSim.m:38-59)
    and f_g0_scalar (private/f_g0_scalar.m:6)."""
    px2 = (wig["pxe"] ** 2).reshape(1, -1, 1, 1)
    py2 = (wig["pye"] ** 2).reshape(-1, 1, 1, 1)
    pz2k = (wig["pze"] ** 2 - sim.k02e).reshape(1, 1, -1, 1)
    E = f_wiggle(E, *_ramps(wig))
    E = _fftn(E, workers=workers)
    E = E / (px2 + py2 + pz2k)
    E = _fftn(E, inverse=True, workers=workers)
    return f_wiggle(E, *_ramps(wig, conj=True))


def f_g0_vector(fEx, fEy, fEz, px, py, pz, k0e):
    """private/f_g0_vector.m:4-8 -- dyadic Green's function (pre-scaled)."""
    div = (px * fEx + py * fEy + pz * fEz) / k0e
    g0 = 1 / (px ** 2 + py ** 2 + pz ** 2 - k0e)
    return g0 * (fEx - px * div), g0 * (fEy - py * div), g0 * (fEz - pz * div)


def propagate_vector(sim: Sim, E, wig, workers=None):
    """WaveSimVector.propagate (WaveSimVector.m:19-51)."""
    E = f_wiggle(E, *_ramps(wig))
    F = _fftn(E, workers=workers)
    px = wig["pxe"].reshape(1, -1, 1)
    py = wig["pye"].reshape(-1, 1, 1)
    pz = wig["pze"].reshape(1, 1, -1)
    fx, fy, fz = f_g0_vector(F[..., 0], F[..., 1], F[..., 2], px, py, pz, sim.k02e)
    E = _fftn(np.stack([fx, fy, fz], axis=3), inverse=True, workers=workers)
    return f_wiggle(E, *_ramps(wig, conj=True))


def propagate(sim: Sim, E, wig, workers=None):
    if sim.vector:
        return propagate_vector(sim, E, wig, workers)
    return propagate_scalar(sim, E, wig, workers)


def crop(sim: Sim, E):
    """Simulation.crop_field (Simulation.m:254-261)."""
    r = sim.roi
    return E[r[0, 0] - 1:r[1, 0], r[0, 1] - 1:r[1, 1], r[0, 2] - 1:r[1, 2], r[0, 3] - 1:r[1, 3]]


def energy(E) -> float:
    """Simulation.energy (Simulation.m:265-273), accumulated in float64."""
    E = np.asarray(E)
    return float(np.sum(E.real.astype(np.float64) ** 2 + E.imag.astype(np.float64) ** 2))


# ----------------------------------------------------------------------------
# the loop (WaveSimBase.m:157-197) and the stopping rule (Simulation.m:193-217)
# ----------------------------------------------------------------------------
@dataclass
class State:
    it: int = 1
    max_iterations: float = math.inf
    diff_energy: list = field(default_factory=list)
    last_step_energy: float = math.inf
    calculate_energy: bool = True
    has_next: bool = True
    converged: bool = False
    source_energy: float = 0.0
    E_primed: Optional[np.ndarray] = None   # full-grid E' (what the MEX returns)
    energy_every: list = field(default_factory=list)  # |crop dE'|^2 at every iteration


def next_state(sim: Sim, state: State, can_terminate: bool,
               callback: Optional[Callable] = None, ref_energy: Optional[float] = None) -> None:
    """Simulation.next (Simulation.m:193-217).  ``ref_energy`` replaces the
    denominator diff_energy(1) (no reference counterpart: a start-independent
    relative residual for warm starts)."""
    state.diff_energy.append(state.last_step_energy)
    de = state.diff_energy
    den = ref_energy if ref_energy is not None and ref_energy > 0 else de[0]
    if can_terminate and de[-1] / den < sim.energy_threshold:
        state.has_next = False
        state.converged = True
    elif can_terminate and state.it >= state.max_iterations:
        state.has_next = False
        state.converged = False
    else:
        state.has_next = True
    state.calculate_energy = ((state.it - 1) % sim.energy_calculation_interval) == 0
    if callback is not None and (state.it % sim.callback_interval == 0 or not state.has_next):
        callback(sim, state)
    state.it += 1


def run(sim: Sim, source: Source, workers=None, callback: Optional[Callable] = None,
        max_iterations: Optional[int] = None, record_every: bool = False,
        E0: Optional[np.ndarray] = None, ref_energy: Optional[float] = None):
    """Simulation.exec (Simulation.m:138-190) + WaveSimBase.run_algorithm
    MATLAB branch (WaveSimBase.m:157-197).  Returns (E_roi, state); the
    full-grid primed field E' is ``state.E_primed``.

    Warm start (no reference counterpart): with ``E0`` (unprimed, full grid)
    and W0 = gamma E0, E' starts at W0 and during the first Nw iterations the
    correction W0/Nw enters the increment recurrence of WaveSimBase.m:96-124
    through each wiggle in turn, next to the reference's source portions:
        X = dE' + W0/Nw,  dE' <- f_mix(X, propagate(X + i/eps/Nw S)) - W0/Nw.
    With E0 = 0 this is exactly the zero-start loop."""
    src = prepare_source(sim, source)
    state = State()
    state.source_energy = src.energy()
    if state.source_energy == 0:
        return None, state
    state.max_iterations = (math.ceil(sim.max_cycles * sim.iterations_per_cycle)
                            if math.isfinite(sim.max_cycles) else math.inf)
    if max_iterations is not None:
        state.max_iterations = max_iterations
    cdt = sim.cdtype
    N = sim.N
    E = np.zeros(N, cdt)
    dE = np.zeros(N, cdt)
    nw = len(sim.wiggles)
    if sim.single:
        A = np.complex64(1j) / np.float32(sim.epsilon) / np.float32(nw)
    else:
        A = 1j / float(sim.epsilon) / nw
    g4 = sim.gamma[..., None]
    warm = E0 is not None
    if warm:
        W = (g4 * np.asarray(E0).reshape(N)).astype(cdt)
        E = W.copy()
        Wn = (W * (np.float32(1) / np.float32(nw) if sim.single else 1.0 / nw)).astype(cdt)
    while state.has_next:
        wig = sim.wiggles[(state.it - 1) % nw]
        if warm and state.it <= nw:
            X = dE + Wn
            Etmp = propagate(sim, src.add_to(X.copy(), A), wig, workers)
            dE = f_mix(X, Etmp, g4, sim.single) - Wn
            E = E + dE
        else:
            if state.it <= nw:
                Etmp = src.add_to(dE.copy(), A)
            else:
                Etmp = dE
            Etmp = propagate(sim, Etmp, wig, workers)
            dE = f_mix(dE, Etmp, g4, sim.single)
            E = E + dE
        if record_every:
            state.energy_every.append(energy(crop(sim, dE)))
        if state.calculate_energy:
            state.last_step_energy = energy(crop(sim, dE))
        next_state(sim, state, state.it % nw == 0, callback, ref_energy)
    state.E_primed = E
    with np.errstate(divide="ignore", invalid="ignore"):
        Ef = E / g4
    return crop(sim, Ef), state


def stop_iteration(diff_like: Sequence[float], nw: int, threshold: float,
                   interval: int, max_iterations=math.inf):
    """Replays the stopping rule on a per-iteration energy sequence
    ``diff_like[k]`` (energy of iteration k+1) and returns (iterations, converged,
    diff_energy) exactly as the loop above would (Simulation.m:193-217)."""
    it, calc, last = 1, True, math.inf
    de = []
    while True:
        if calc:
            last = diff_like[it - 1]
        de.append(last)
        can = it % nw == 0
        if can and de[-1] / de[0] < threshold:
            return it, True, de
        if can and it >= max_iterations:
            return it, False, de
        calc = ((it - 1) % interval) == 0
        it += 1
