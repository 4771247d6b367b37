"""numpy restatement of the warm-start preprocessing kernels (TEST INFRASTRUCTURE ONLY).

No reference counterpart: these check the project's wfno_wkb_prior,
wfno_features and wfno_compose_prior kernels.  Layout [z][y][x] (x fastest)
for ROI arrays, [y][x][z] (MATLAB index order) for the solver grid.

Pinned by closed forms (tests/test_oracle_warmstart.py): a homogeneous medium
gives the background itself, constant contrast a linear phase, a slab the
path integral per ray; block means by explicit slices; zero envelope gives the
prior.
"""
from __future__ import annotations

import numpy as np


def wkb_prior(n_zyx, u_roi, k0h, n_bg):
    """phi(z) = k0 h (sum_{z'<z}(n - n_bg) + (n(z) - n_bg)/2); P = u(z) exp(i phi)."""
    d = n_zyx.astype(np.float64) - n_bg
    phase = k0h * (np.cumsum(d, axis=0) - 0.5 * d)
    P = u_roi[:, None, None] * np.exp(1j * phase)
    return phase, P


def features(n_zyx, phase, f, n_bg):
    def pool(a):
        Z, Y, X = a.shape
        return a.reshape(Z // f, f, Y // f, f, X // f, f).mean(axis=(1, 3, 5))
    L = n_zyx.shape[0] // f
    c = (np.arange(L) + 0.5) / L - 0.5
    cz, cy, cx = np.meshgrid(c, c, c, indexing="ij")
    return np.stack([pool(n_zyx.astype(np.float64) ** 2 - n_bg ** 2), pool(np.cos(phase)), pool(np.sin(phase)),
                     cz, cy, cx])


def compose(grid, roi_lo, roi_n, f, P, A, u, foot):
    """E0 on the solver grid [y][x][z]: P (1 + A) in the ROI, u(z) on the
    source footprint elsewhere (grid/roi in [y x z] order)."""
    Ny, Nx, Nz = grid
    E = np.zeros((Ny, Nx, Nz), complex)
    E[foot[0]:foot[1], foot[2]:foot[3], :] = u[None, None, :]
    y0, x0, z0 = roi_lo
    ny, nx, nz = roi_n
    Aup = (A[0] + 1j * A[1]).repeat(f, 0).repeat(f, 1).repeat(f, 2)      # [z][y][x]
    E[y0:y0 + ny, x0:x0 + nx, z0:z0 + nz] = (P * (1 + Aup)).transpose(1, 2, 0)
    return E
