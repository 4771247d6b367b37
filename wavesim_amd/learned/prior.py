"""Closed-form background field used as the physics input of the warm start.

For a plane source (a y-x plane of constant amplitude) in a homogeneous
medium of index n_bg, the field is the 1-D band-limited Green's function of
the reference utility homogeneous_medium_analytic_solution.m:26-35 along z:

    u(z) = i h/(2k) e^{ik|z|} - h/(4 pi k) [ e^{ik|z|}(E1(i(k-pi/h)|z|) - E1(i(k+pi/h)|z|))
                                           - e^{-ik|z|}(-E1(-i(k-pi/h)|z|) + E1(-i(k+pi/h)|z|)) ]

with E1 the exponential integral (MATLAB expint).  The field is placed on the
source plane's transverse footprint and taken as zero elsewhere.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.special as sp


def green_1d(k: float, h: float, z) -> np.ndarray:
    z = np.abs(np.asarray(z, dtype=np.float64))
    phi = k * z
    with np.errstate(all="ignore"):
        u = (1j * h / (2 * k) * np.exp(1j * phi)
             - h / (4 * math.pi * k) * (
                 np.exp(1j * phi) * (sp.exp1(1j * (k - math.pi / h) * z) - sp.exp1(1j * (k + math.pi / h) * z))
                 - np.exp(-1j * phi) * (-sp.exp1(-1j * (k - math.pi / h) * z) + sp.exp1(-1j * (k + math.pi / h) * z))))
    u0 = 1j * h / (2 * k) * (1 + 2j * math.atanh(h * k / math.pi) / math.pi)
    return np.where(np.abs(k * z) < 1e-10, u0, u)


def plane_wave_background(N, dx: float, k: float, z_plane: int, y_range, x_range, amplitude=1.0):
    """Unprimed background field on the full grid N=[Ny,Nx,Nz,1] for a source
    plane at 0-based grid index z_plane covering rows y_range and columns
    x_range (python slices)."""
    Ny, Nx, Nz = N[:3]
    u = green_1d(k, dx, (np.arange(Nz) - z_plane) * dx) * amplitude
    E = np.zeros((Ny, Nx, Nz, 1), np.complex64)
    E[y_range, x_range, :, 0] = u.astype(np.complex64)[None, None, :]
    return E
