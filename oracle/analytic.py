"""Closed-form solutions used to pin the oracle (TEST INFRASTRUCTURE ONLY).

* ``homogeneous_1d`` restates utilities/homogeneous_medium_analytic_solution.m:1-36
  (the band-limited 1-D Green's function of the reference), with MATLAB's
  ``expint`` = E1 = ``scipy.special.exp1`` (principal branch for complex z).
* ``periodic_homogeneous`` is the exact discrete solution of the periodic
  problem  (|p|^2 - k0^2 n^2) E(p) = S(p)  that a converged MBS run with no
  boundaries (so no filters and one wiggle) must reproduce.
* ``helmholtz_residual`` evaluates the spectral Helmholtz operator
  (-laplace - k0^2 n^2) E - S on a periodic grid.
* ``hankel_2d`` is the 2-D free-space Green's function (i h^2/4) H0(k r) for a
  unit pixel source of area h^2.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.special as sp


def homogeneous_1d(k: float, h: float, x) -> np.ndarray:
    """utilities/homogeneous_medium_analytic_solution.m:26-35."""
    x = np.abs(np.asarray(x, dtype=np.float64))
    phi = k * x
    with np.errstate(all="ignore"):
        sol = (1.0j * h / (2 * k) * np.exp(1.0j * phi)
               - h / (4 * math.pi * k) * (
                   np.exp(1.0j * phi) * (sp.exp1(1.0j * (k - math.pi / h) * x)
                                         - sp.exp1(1.0j * (k + math.pi / h) * x))
                   - np.exp(-1.0j * phi) * (-sp.exp1(-1.0j * (k - math.pi / h) * x)
                                            + sp.exp1(-1.0j * (k + math.pi / h) * x))))
    small = np.abs(k * x) < 1e-10
    sol = np.where(small, 1.0j * h / (2 * k) * (1 + 2j * math.atanh(h * k / math.pi) / math.pi), sol)
    return sol


def fourier_coords(N, dx):
    """Angular wavenumbers p = 2 pi/(dx N) * symrange(N) for each of [y, x, z]
    (SimulationGrid.m:45-50), as broadcastable arrays."""
    out = []
    for d, n in enumerate(N):
        p = 2 * math.pi / (dx * n) * np.fft.fftshift(np.arange(n) - n // 2)
        shape = [1, 1, 1]
        shape[d] = n
        out.append(p.reshape(shape))
    return out


def periodic_homogeneous(S: np.ndarray, k0: float, n: complex, dx: float) -> np.ndarray:
    """Exact periodic solution of  laplace E + k0^2 n^2 E = -S  (spectral
    Laplacian): E = ifftn( fftn(S) / (|p|^2 - k0^2 n^2) ).  S has shape (Ny,Nx,Nz)."""
    py, px, pz = fourier_coords(S.shape, dx)
    den = px ** 2 + py ** 2 + pz ** 2 - (k0 * n) ** 2
    return np.fft.ifftn(np.fft.fftn(S.astype(np.complex128)) / den)


def helmholtz_residual(E: np.ndarray, S: np.ndarray, k0: float, n_map: np.ndarray,
                       dx: float) -> np.ndarray:
    """(-laplace - k0^2 n^2) E - S with the spectral Laplacian on a periodic grid."""
    py, px, pz = fourier_coords(E.shape, dx)
    lapE = np.fft.ifftn(-(px ** 2 + py ** 2 + pz ** 2) * np.fft.fftn(E.astype(np.complex128)))
    return -lapE - (k0 ** 2) * (n_map.astype(np.complex128) ** 2) * E - S


def hankel_2d(k: float, h: float, r) -> np.ndarray:
    """(i h^2 / 4) H0^(1)(k r): field of a unit pixel source in 2-D."""
    return 1.0j * h * h / 4 * sp.hankel1(0, k * np.asarray(r, dtype=np.float64))
