"""Synthetic workloads of the benchmark configurations (seeded, no data files).

* ``disordered_3d``: n = 1 + 0.5 u, u = min-max rescaled U(0,1) noise smoothed
  by a Gaussian (sigma 2 px, periodic), 4 px per wavelength, ARL boundary,
  point source at the centre (examples/disordered_medium_3D.m:44 pattern).
* ``disordered_2d``: n = 1 + 0.5 U(0,1) smoothed (sigma 2 px) + 0.01i,
  plane-wave source ones(Ny,1) at [1,1] (examples/pulse2d.m:34 pattern).
* ``homogeneous_2d``: n = 1 with a unit point source at the centre
  (examples/homogeneous_medium_2D.m:25-26 pattern).
* ``sphere_vector_3d``: one dielectric sphere (n 1.6 in 1.33), x-polarised
  plane wave (examples/mie_sphere/run_mie_sphere_simulation.m:14-21 pattern).
"""
from __future__ import annotations

import numpy as np
import scipy.ndimage as ndi

from .simulation import Source


def _smoothed_uniform(shape, seed, sigma=2.0):
    rng = np.random.default_rng(seed)
    u = ndi.gaussian_filter(rng.random(shape, dtype=np.float32), sigma, mode="wrap")
    return (u - u.min()) / (u.max() - u.min())


def disordered_3d(n=256, boundary=32, seed=0, residual=1e-3):
    u = _smoothed_uniform((n, n, n), seed)
    nmap = (1.0 + 0.5 * u).astype(np.float32)
    opt = dict(pixel_size=0.25, boundary_widths=[boundary] * 3, energy_threshold=residual ** 2,
               callback_interval=1e9)
    return nmap, opt, Source(1.0, [n // 2, n // 2, n // 2])


def disordered_2d(n=1024, boundary=64, seed=0, residual=1e-3):
    rng = np.random.default_rng(seed)
    u = ndi.gaussian_filter(rng.random((n, n)), 2.0, mode="wrap")
    nmap = 1.0 + 0.5 * u + 0.01j
    opt = dict(pixel_size=0.25, boundary_widths=[boundary, boundary], energy_threshold=residual ** 2,
               callback_interval=1e9)
    return nmap, opt, Source(np.ones(n), [1, 1])


def homogeneous_2d(n=256, boundary=32, residual=1e-3):
    opt = dict(pixel_size=0.25, boundary_widths=[boundary, boundary], energy_threshold=residual ** 2,
               callback_interval=1e9)
    return np.ones((n, n)), opt, Source(1.0, [n // 2, n // 2])


def sphere_vector_3d(n=128, boundary=32, radius=20, residual=1e-3):
    c = (n - 1) / 2
    g = np.arange(n) - c
    r2 = g[:, None, None] ** 2 + g[None, :, None] ** 2 + g[None, None, :] ** 2
    nmap = np.where(r2 <= radius ** 2, 1.6, 1.33)
    opt = dict(pixel_size=0.25, boundary_widths=[boundary] * 3, energy_threshold=residual ** 2,
               callback_interval=1e9)
    return nmap, opt, Source(np.ones((n, n, 1, 1)), [1, 1, 1, 1])


def spheres_3d(n=128, boundary=32, seed=0, residual=1e-3, n_bg=1.33):
    """1-8 seeded random spheres (radius 4-16 px, n 1.2-1.6) in a background
    n_bg, plane-wave source: the plane z = 1 of the ROI (propagation along z)."""
    rng = np.random.default_rng(seed)
    nmap = np.full((n, n, n), n_bg, dtype=np.float32)
    g = np.arange(n, dtype=np.float32)
    for _ in range(int(rng.integers(1, 9))):
        r = rng.uniform(4, 16)
        c = rng.uniform(r, n - r, size=3)
        ns = rng.uniform(1.2, 1.6)
        d2 = ((g[:, None, None] - c[0]) ** 2 + (g[None, :, None] - c[1]) ** 2 +
              (g[None, None, :] - c[2]) ** 2)
        nmap[d2 <= r * r] = ns
    opt = dict(pixel_size=0.25, boundary_widths=[boundary] * 3, energy_threshold=residual ** 2,
               callback_interval=1e9)
    return nmap, opt, Source(np.ones((n, n, 1)), [1, 1, 1])


RAGGED_SIZES = (256, 384, 512, 768, 1024, 2048)  # interiors; 32 px boundary, grids padded to FFT sizes


def ragged_2d(count=32, boundary=32, residual=1e-3, sizes=RAGGED_SIZES):
    """``count`` independent 2-D disordered problems of ragged sizes (cycling
    through ``sizes``), each its own seeded medium n = 1 + 0.5 u + 0.01i
    (Gaussian-smoothed, sigma 2 px) with a plane-wave source along x: the
    ``disordered_2d`` pattern at several sizes.  Grids 768 + 64 and
    1024 + 64 are padded to 840 and 1100 (SimulationGrid.m:64-80)."""
    out = []
    for i in range(count):
        n = sizes[i % len(sizes)]
        rng = np.random.default_rng(1000 + i)
        u = ndi.gaussian_filter(rng.random((n, n)), 2.0, mode="wrap")
        nmap = 1.0 + 0.5 * u + 0.01j
        opt = dict(pixel_size=0.25, boundary_widths=[boundary, boundary], energy_threshold=residual ** 2,
                   callback_interval=1e9)
        out.append((nmap, opt, Source(np.ones(n), [1, 1])))
    return out


WORKLOADS = {
    "disordered_3d_256": lambda: disordered_3d(256, 32),
    # the medium of the 64-source batch (192^3 + 32 px boundary: a 256^3 grid)
    "disordered_3d_192": lambda: disordered_3d(192, 32),
    "disordered_2d_1024": lambda: disordered_2d(1024, 64),
    "homogeneous_2d_256": lambda: homogeneous_2d(256, 32),
    "sphere_vector_3d_128": lambda: sphere_vector_3d(128, 32),
    "spheres_3d_128": lambda: spheres_3d(128, 32),
}
