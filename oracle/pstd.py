"""numpy restatement of the reference's PSTD comparison solver (PSTD.m).

TEST INFRASTRUCTURE ONLY: the checker for the HIP PSTD path, never the product.

The reference's pseudospectral time-domain solver steps the scalar wave
equation  nabla^2 E - e_r d^2E/dt^2 = -source  with a centred difference in
time and the spectral Laplacian ifftn(-|p|^2 fftn(E)) (PSTD.m:98-111), under a
CW source whose amplitude ramps up over 100 optical cycles (PSTD.m:130-137),
until the change per step, measured against the expected phase advance of the
source, falls below the energy threshold (PSTD.m:113-121, Simulation.m:193-217).

Conventions as ``oracle.mbs``: arrays [y, x, z, pol], positions and ROIs
1-based.  ``single=True`` (the reference default) keeps the fields in
complex64; the coefficient arrays c1, c2, c3 and the Laplacian are computed in
double (PSTD.m:57-68) and enter single-precision products, as MATLAB's mixed
single/double arithmetic does.

Pins (tests/test_oracle_pstd.py): the construction constants against hand
computations; the Laplacian against -|p|^2 on grid plane waves; one step of the
recurrence against the closed form cos(k theta) of a lossless Fourier mode
(cos theta = 1 - dt^2 |p|^2 / (2 e_r)); the converged field of a homogeneous
absorbing periodic medium against the closed-form time-harmonic solution of
the recurrence, X / z with X = c3 S / (1 - (c2 - c3 |p|^2) / z - c1 / z^2),
z = exp(-2 pi i / iterations_per_cycle); the stopping rule against hand traces.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import mbs


@dataclass
class PSTDSim:
    N: List[int]               # [Ny, Nx, Nz, 1]
    roi: np.ndarray            # 2x4, 1-based inclusive
    medium: mbs.MediumData
    dimensions: int
    dtmax: float
    dt: float
    iterations_per_cycle: float
    omega: float
    c1: np.ndarray             # (Ny, Nx, Nz) float64
    c2: np.ndarray
    c3: np.ndarray
    koperator: np.ndarray      # (Ny, Nx, Nz) float64, -|p|^2
    energy_threshold: float
    energy_calculation_interval: int
    callback_interval: float
    max_cycles: float          # the option + 100 (PSTD.m:69)
    single: bool

    @property
    def cdtype(self):
        return np.complex64 if self.single else np.complex128

    @property
    def rdtype(self):
        return np.float32 if self.single else np.float64


def make_pstd(refractive_index, opt: dict) -> PSTDSim:
    """PSTD constructor (PSTD.m:22-77) on the Simulation constructor
    (Simulation.m:80-136) and Medium (Medium.m:15-72)."""
    # PSTD.m:32-34: boundary_type must be given and boundary_strength must not
    if "boundary_type" not in opt or "boundary_strength" in opt:
        raise ValueError("options field boundary_type and boundary_strength needs to be specified")
    opt = dict(opt)
    opt["ACC"] = False                                            # PSTD.m:35
    single = bool(opt.get("single_precision", True))
    lam = float(opt.get("lambda", 1.0))
    callback_interval = opt.get("callback_interval", 50)
    eci = min(opt.get("energy_calculation_interval", 8), callback_interval)   # Simulation.m:135
    med = mbs.medium(refractive_index, opt)
    grid = med.grid
    N = list(grid.N) + [1]
    default_roi = np.array([[b + 1 for b in med.Bl], [grid.N[d] - med.Br[d] for d in range(3)]])
    if opt.get("roi") is None:                                    # Simulation.m:118-124
        roi = np.array([mbs.make4(default_roi[0]), mbs.make4(default_roi[1])])
    else:
        r = np.asarray(opt["roi"])
        roi = np.array([mbs.make4(r[0]), mbs.make4(r[1])])
        roi[:, :3] = roi[:, :3] + default_roi[0] - 1
    dimensions = 2 if min(grid.N) == 1 else 3                     # PSTD.m:45-49
    dtmax = 2 / math.sqrt(dimensions) / math.pi * grid.dx * math.sqrt(med.e_r_min)   # PSTD.m:51
    dt_relative = float(opt.get("dt_relative", 0.95))
    dt = dt_relative * dtmax                                      # PSTD.m:52
    ipc = lam / dt                                                # PSTD.m:53
    omega = 2 * math.pi / lam                                     # PSTD.m:56
    e_r = med.e_r
    c2dt = dt ** 2 / e_r.real                                     # PSTD.m:57
    sdt = dt * omega * e_r.imag / e_r.real                        # PSTD.m:58
    c1 = (sdt - 2) / (sdt + 2)                                    # PSTD.m:60
    c2 = 4 / (sdt + 2)                                            # PSTD.m:61
    c3 = 2 * c2dt / (sdt + 2)                                     # PSTD.m:62
    px = grid.px_range.reshape(1, -1, 1)                          # PSTD.m:65-68
    py = grid.py_range.reshape(-1, 1, 1)
    pz = grid.pz_range.reshape(1, 1, -1)
    kop = -px ** 2 - py ** 2 - pz ** 2
    kop = np.broadcast_to(kop, tuple(grid.N)).copy()
    return PSTDSim(N=N, roi=roi, medium=med, dimensions=dimensions, dtmax=dtmax, dt=dt,
                   iterations_per_cycle=ipc, omega=omega, c1=c1, c2=c2, c3=c3, koperator=kop,
                   energy_threshold=float(opt.get("energy_threshold", 1e-2)),
                   energy_calculation_interval=int(eci), callback_interval=callback_interval,
                   max_cycles=float(opt.get("max_cycles", math.inf)) + 100, single=single)


def default_source(ncycle: float) -> complex:
    """PSTD.default_source (PSTD.m:130-137): exp(-2 pi i N) times a raised
    cosine that reaches 1 after 100 cycles."""
    A = np.exp(-2.0j * math.pi * ncycle)
    return complex(A * (0.5 - 0.5 * math.cos(math.pi * min(ncycle / 100.0, 1.0))))


def laplacian(sim: PSTDSim, E):
    """ifftn(koperator .* fftn(E)) (PSTD.m:108), over the three grid axes."""
    k = sim.koperator[..., None].astype(sim.rdtype) if sim.single else sim.koperator[..., None]
    F = mbs._fftn(np.asarray(E))
    return mbs._fftn((k * F).astype(sim.cdtype), inverse=True).astype(sim.cdtype)


def step(sim: PSTDSim, E, E_prev, src: mbs.Source, A: complex):
    """One update (PSTD.m:108-111): E_next = c2 E + c1 E_prev + c3 (lap E + A S)."""
    Etmp = laplacian(sim, E)
    if src is not None:
        a = np.complex64(A) if sim.single else complex(A)
        Etmp = src.add_to(Etmp, a)
    if sim.single:
        c1, c2, c3 = (c[..., None].astype(np.float32) for c in (sim.c1, sim.c2, sim.c3))
    else:
        c1, c2, c3 = (c[..., None] for c in (sim.c1, sim.c2, sim.c3))
    return ((c2 * E + c1 * E_prev) + c3 * Etmp).astype(sim.cdtype)


@dataclass
class PSTDState:
    it: int = 1
    max_iterations: float = math.inf
    diff_energy: list = field(default_factory=list)
    last_step_energy: float = math.inf
    calculate_energy: bool = True
    has_next: bool = True
    converged: bool = False
    source_energy: float = 0.0
    E_full: Optional[np.ndarray] = None   # full-grid state.E after the last step (before the phase)


def _next(sim: PSTDSim, state: PSTDState, can_terminate: bool) -> None:
    """Simulation.next (Simulation.m:193-217) with PSTD's can_terminate."""
    state.diff_energy.append(state.last_step_energy)
    de = state.diff_energy
    if can_terminate and de[-1] / de[0] < sim.energy_threshold:
        state.has_next = False
        state.converged = True
    elif can_terminate and state.it >= state.max_iterations:
        state.has_next = False
        state.converged = False
    else:
        state.has_next = True
    state.calculate_energy = ((state.it - 1) % sim.energy_calculation_interval) == 0
    state.it += 1


def run(sim: PSTDSim, source: mbs.Source, max_iterations: Optional[int] = None):
    """Simulation.exec (Simulation.m:138-190) + PSTD.run_algorithm
    (PSTD.m:79-127).  Returns (E on the ROI, phase-compensated; state)."""
    src = source.cast(sim.rdtype).shift(sim.roi[0]).crop(np.array([[1, 1, 1, 1], sim.N]))
    state = PSTDState()
    state.source_energy = src.energy()
    if state.source_energy == 0:
        return None, state
    state.max_iterations = (math.ceil(sim.max_cycles * sim.iterations_per_cycle)
                            if math.isfinite(sim.max_cycles) else math.inf)
    if max_iterations is not None:
        state.max_iterations = max_iterations
    E = np.zeros(sim.N, sim.cdtype)                               # PSTD.m:81-83
    E_prev = np.zeros(sim.N, sim.cdtype)
    A = 1.0 + 0.0j
    while state.has_next:
        ncycle = state.it / sim.iterations_per_cycle              # PSTD.m:94
        A_prev = A
        A = default_source(ncycle)                                # PSTD.m:95-96
        E_next = step(sim, E, E_prev, src, A)                     # PSTD.m:108-111
        if state.calculate_energy:                                # PSTD.m:113-116
            ps = np.exp(1.0j * (np.angle(A) - np.angle(A_prev)))
            p = np.complex64(ps) if sim.single else ps
            d = (E_next - E * p).astype(sim.cdtype)
            state.last_step_energy = mbs.energy(mbs.crop(sim, d)) / abs(A) ** 2
        E_prev = E                                                # PSTD.m:118-119
        E = E_next
        can_terminate = A.real > 0.9                              # PSTD.m:120 (real part, as MATLAB compares)
        _next(sim, state, can_terminate)
    state.E_full = E
    A_next = default_source(state.it / sim.iterations_per_cycle)  # PSTD.m:125-126
    ph = np.exp(-1.0j * np.angle(A_next))
    Eout = mbs.crop(sim, E) * (np.complex64(ph) if sim.single else ph)
    return Eout.astype(sim.cdtype), state


def steady_state_periodic(sim: PSTDSim, S_dense: np.ndarray) -> np.ndarray:
    """Closed-form time-harmonic solution of the recurrence for a homogeneous
    medium on a periodic grid (c1, c2, c3 constant), full grid, phase
    compensated as run() returns it: with A_k = z^k, z = exp(-2 pi i / ipc),
    E_k = X z^k per Fourier mode, X = c3 S / (1 - (c2 - c3 |p|^2)/z - c1/z^2),
    and the output E_K exp(-i angle(A_{K+1})) = X / z."""
    z = np.exp(-2j * math.pi / sim.iterations_per_cycle)
    c1, c2, c3 = float(sim.c1.flat[0]), float(sim.c2.flat[0]), float(sim.c3.flat[0])
    p2 = -sim.koperator
    F = np.fft.fftn(S_dense[..., 0].astype(np.complex128), axes=(0, 1, 2))
    X = c3 * F / (1 - (c2 - c3 * p2) / z - c1 / z ** 2)
    return (np.fft.ifftn(X, axes=(0, 1, 2)) / z)[..., None]
