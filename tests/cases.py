"""Seeded problem builders shared by the oracle tests and the GPU parity tests.

Each builder returns (refractive_index, options, [(value, position), ...]) in
the reference's conventions ([y, x, z(, pol)] order, 1-based positions
relative to the ROI); ``oracle_source`` / ``gpu_source`` turn the source list
into the oracle's and the package's Source objects.
"""
from __future__ import annotations

import numpy as np
import scipy.ndimage as ndi


def smooth_disorder(shape, seed, lo=1.0, span=0.5, sigma=2.0, absorption=0.0):
    """n = lo + span * u, u = min-max rescaled Gaussian-smoothed U(0,1) noise."""
    rng = np.random.default_rng(seed)
    u = ndi.gaussian_filter(rng.random(shape), sigma, mode="wrap")
    u = (u - u.min()) / (u.max() - u.min())
    return lo + span * u + 1j * absorption


def sphere(shape, radius, n_in, n_out, center=None):
    c = [(s - 1) / 2 for s in shape] if center is None else center
    grids = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    r2 = sum((g - ci) ** 2 for g, ci in zip(grids, c))
    return np.where(r2 <= radius ** 2, n_in, n_out).astype(np.complex128)


def base_opt(bw, thr=1e-6, **kw):
    o = dict(pixel_size=0.25, boundary_widths=bw, energy_threshold=thr, callback_interval=50)
    o.update(kw)
    return o


def case_2d_homogeneous(n=48, b=8, thr=1e-6):
    return np.ones((n, n)), base_opt([b, b], thr), [(1.0, [n // 2, n // 2])]


def case_2d_disordered(ny=40, nx=56, b=8, seed=1, thr=1e-6):
    nmap = smooth_disorder((ny, nx), seed, absorption=0.01)
    return nmap, base_opt([b, b], thr), [(np.ones(ny), [1, 1])]


def case_3d_disordered(shape=(20, 24, 16), b=4, seed=2, thr=1e-5):
    nmap = smooth_disorder(shape, seed)
    c = [s // 2 for s in shape]
    return nmap, base_opt([b, b, b], thr), [(1.0, c)]


def case_3d_vector_sphere(n=16, b=4, thr=1e-5):
    nmap = sphere((n, n, n), n / 4, 1.6, 1.33)
    src = np.ones((n, n, 1, 1), dtype=np.complex128)  # x-polarised plane
    return nmap, base_opt([b, b, b], thr), [(src, [1, 1, 1, 1])]


def case_1d(n=64, b=16, thr=1e-8):
    return np.ones((n, 1)), base_opt([b, 0], thr), [(1.0, [n // 2, 1])]


def case_periodic_2d(ny=24, nx=30, seed=3, thr=1e-10):
    nmap = smooth_disorder((ny, nx), seed, absorption=0.05)
    return nmap, base_opt([0, 0], thr), [(1.0, [5, 7])]


def oracle_source(src_list):
    from oracle import mbs
    s = mbs.Source()
    for v, p in src_list:
        s = s + mbs.Source.point(v, p)
    return s


def gpu_source(src_list):
    from wavesim_amd import Source
    s = Source()
    for v, p in src_list:
        s = s + Source(v, p)
    return s


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def divergence_stop(energy_every, nw, ref=0.0):
    """The failure exit of the library's stop rule, restated over a history of
    ROI increment energies (one per iteration): the first iteration whose
    energy is not finite, or, at a multiple of Nw after the second cycle,
    exceeds the energies one and two cycles earlier (which increase) and 100
    times max(first energy, ref).  None when no iteration qualifies."""
    e = np.asarray(energy_every, dtype=np.float64)
    for k in range(len(e)):
        it = k + 1
        if not np.isfinite(e[k]):
            return it
        if it % nw == 0 and it > 2 * nw:
            e1, e2 = e[k - nw], e[k - 2 * nw]
            if e[k] > e1 and e1 > e2 and e[k] > 100.0 * max(e[0], ref):
                return it
    return None
