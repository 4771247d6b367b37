"""Mie theory: the field of a plane wave scattered by a homogeneous sphere.

TEST INFRASTRUCTURE ONLY (the checker of the vector solver's accuracy; the
product never imports it).

The reference computes this analytic solution for its Mie-sphere example
(examples/mie_sphere/mie_sphere_analytical_solution.m:31) with the
third-party package MatScat (J. Schaefer, 2012, bundled as
examples/mie_sphere/MatScat: calcmie_nf.m:63-82, util/nfmie.m:40-150,
expcoeff/expcoeff_mie.m:20-33, expcoeff/expcoeff_mie_int.m:17-30).  This
file restates the published method MatScat follows -- Bohren & Huffman,
"Absorption and Scattering of Light by Small Particles" (Wiley 1983),
sections 4.1-4.4 -- with the same conventions, so that its output means what
MatScat's default output means:

* incident field E_i = exp(i k z) x_hat (time dependence exp(-i w t), the
  convention of the solver: waves exp(+i k z) travel towards +z);
* k = 2 pi n_medium / lambda, size parameter x = k a, relative index m;
* outside the sphere the *scattered* field sum_n E_n (i a_n N_e1n - b_n M_o1n)
  with spherical Hankel functions h_n^(1) (BH 4.45);
* inside, the internal field sum_n E_n (c_n M_o1n - i d_n N_e1n) with j_n
  (BH 4.46) minus the incident field, i.e. the total minus the incident field
  everywhere (MatScat's TotalField = false);
* E_n = i^n (2n+1) / (n(n+1)), truncation after n_max = ceil(x + 4 x^(1/3)
  + 2) orders (Wiscombe's criterion, BH appendix A);
* Cartesian components (x, y, z) of the field.

Pins (tests/test_oracle_mie.py): energy conservation of a lossless sphere
(Q_ext = Q_sca, BH 4.61-4.62), continuity of tangential E and of normal D
across the surface, the Rayleigh-dipole near field for x << 1, and the
electrostatic internal field 3/(m^2+2) at the centre of a small sphere.
"""
from __future__ import annotations

import numpy as np
from scipy import special


def n_orders(x: float) -> int:
    """Wiscombe's truncation (BH appendix A; expcoeff_mie.m:21)."""
    return int(np.ceil(x + 4.0 * x ** (1.0 / 3.0) + 2.0))


def _psi(n, z):
    """Riccati-Bessel psi_n(z) = z j_n(z) and its derivative."""
    j = special.spherical_jn(n, z)
    dj = special.spherical_jn(n, z, derivative=True)
    return z * j, j + z * dj


def _xi(n, z):
    """Riccati-Bessel xi_n(z) = z h_n^(1)(z) and its derivative."""
    h = special.spherical_jn(n, z) + 1j * special.spherical_yn(n, z)
    dh = special.spherical_jn(n, z, derivative=True) + 1j * special.spherical_yn(n, z, derivative=True)
    return z * h, h + z * dh


def coefficients(x: float, m: complex, nmax: int | None = None):
    """Scattering (a_n, b_n, BH 4.53) and internal (c_n, d_n, BH 4.52)
    coefficients for n = 1..nmax, relative permeability 1."""
    nmax = nmax or n_orders(x)
    n = np.arange(1, nmax + 1)
    px, dpx = _psi(n, x)
    pm, dpm = _psi(n, m * x)
    xx, dxx = _xi(n, x)
    a = (m * pm * dpx - px * dpm) / (m * pm * dxx - xx * dpm)
    b = (pm * dpx - m * px * dpm) / (pm * dxx - m * xx * dpm)
    # (psi xi' - xi psi') = i (Wronskian); kept in its defining form
    w = px * dxx - xx * dpx
    c = m * w / (pm * dxx - m * xx * dpm)
    d = m * w / (m * pm * dxx - xx * dpm)
    return a, b, c, d


def efficiencies(x: float, m: complex):
    """Extinction and scattering efficiencies (BH 4.61, 4.62)."""
    a, b, _, _ = coefficients(x, m)
    n = np.arange(1, a.size + 1)
    qext = 2.0 / x ** 2 * np.sum((2 * n + 1) * np.real(a + b))
    qsca = 2.0 / x ** 2 * np.sum((2 * n + 1) * (np.abs(a) ** 2 + np.abs(b) ** 2))
    return float(qext), float(qsca)


def _angular(mu, nmax):
    """pi_n(cos theta), tau_n(cos theta) for n = 1..nmax (BH 4.47):
    pi_n = (2n-1)/(n-1) mu pi_{n-1} - n/(n-1) pi_{n-2}, tau_n = n mu pi_n -
    (n+1) pi_{n-1}; arrays [points, nmax]."""
    P = mu.size
    pi = np.zeros((P, nmax + 1))
    tau = np.zeros((P, nmax + 1))
    pi[:, 1] = 1.0
    tau[:, 1] = mu
    for n in range(2, nmax + 1):
        pi[:, n] = (2 * n - 1) / (n - 1) * mu * pi[:, n - 1] - n / (n - 1) * pi[:, n - 2]
        tau[:, n] = n * mu * pi[:, n] - (n + 1) * pi[:, n - 1]
    return pi[:, 1:], tau[:, 1:]


def near_field(radius, n_sphere, n_medium, wavelength, x, y, z, chunk=65536):
    """Total minus incident field (scattered outside, internal minus incident
    inside) at Cartesian points x, y, z (same units as radius and
    wavelength), for an x-polarised plane wave of unit amplitude travelling
    along +z; returns an array of shape x.shape + (3,) with the x, y, z
    components (calcmie_nf.m defaults)."""
    x, y, z = (np.asarray(v, dtype=np.float64) for v in np.broadcast_arrays(x, y, z))
    shape = x.shape
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    k = 2 * np.pi * n_medium / wavelength
    m = n_sphere / n_medium
    a_n, b_n, c_n, d_n = coefficients(k * radius, m)
    nmax = a_n.size
    n = np.arange(1, nmax + 1)
    En = (1j ** n) * (2 * n + 1) / (n * (n + 1))
    out = np.empty((x.size, 3), dtype=np.complex128)
    for s in range(0, x.size, chunk):
        e = slice(s, s + chunk)
        out[e] = _field_chunk(x[e], y[e], z[e], radius, k, m, n, En, a_n, b_n, c_n, d_n)
    return out.reshape(shape + (3,))


def _field_chunk(x, y, z, radius, k, m, n, En, a_n, b_n, c_n, d_n):
    r = np.sqrt(x * x + y * y + z * z)
    r = np.maximum(r, 1e-9 * radius)  # the centre: the limit r -> 0 along +z
    theta = np.arccos(np.clip(z / r, -1.0, 1.0))
    phi = np.arctan2(y, x)
    st, ct, sp, cp = np.sin(theta), np.cos(theta), np.sin(phi), np.cos(phi)
    pi, tau = _angular(ct, n.size)
    inside = r < radius
    rho = np.where(inside, m * k * r, k * r)[:, None]
    nn = n[None, :]
    # radial functions: z_n(rho) and (rho z_n)'/rho
    zn = np.empty((x.size, n.size), dtype=np.complex128)
    dz = np.empty((x.size, n.size), dtype=np.complex128)
    if np.any(inside):
        ps, dps = _psi(nn, rho[inside])
        zn[inside], dz[inside] = ps / rho[inside], dps / rho[inside]
    if np.any(~inside):
        xi, dxi = _xi(nn, rho[~inside])
        zn[~inside], dz[~inside] = xi / rho[~inside], dxi / rho[~inside]
    # vector spherical harmonics (BH 4.50): (r, theta, phi) components
    c, s = cp[:, None], sp[:, None]
    M_o = (np.zeros_like(zn), c * pi * zn, -s * tau * zn)
    N_e = (c * nn * (nn + 1) * st[:, None] * pi * zn / rho, c * tau * dz, -s * pi * dz)
    # outside: i a_n N_e - b_n M_o; inside: c_n M_o - i d_n N_e
    A_N = np.where(inside[:, None], -1j * d_n[None, :], 1j * a_n[None, :])
    A_M = np.where(inside[:, None], c_n[None, :], -b_n[None, :])
    Er, Et, Ep = (np.sum(En[None, :] * (A_N * N_e[q] + A_M * M_o[q]), axis=1) for q in range(3))
    out = np.empty((x.size, 3), dtype=np.complex128)
    out[:, 0] = st * cp * Er + ct * cp * Et - sp * Ep
    out[:, 1] = st * sp * Er + ct * sp * Et + cp * Ep
    out[:, 2] = ct * Er - st * Et
    out[inside, 0] -= np.exp(1j * k * z[inside])
    return out


def reference_grid(N, pixel_size, center=(0.0, 0.0, 0.0)):
    """Coordinates of examples/mie_sphere/create_sphere.m:33-37 for a grid
    N = [Ny, Nx, Nz] (kept as written there: the x coordinates are offset by
    N(1)/2 and the y coordinates by N(2)/2 pixels), as [y, x, z] arrays."""
    N = [int(v) for v in N]
    c = np.asarray(N, dtype=np.float64) / 2 * pixel_size + np.asarray(center)
    xr = np.arange(1, N[1] + 1) * pixel_size - c[0]
    yr = np.arange(1, N[0] + 1) * pixel_size - c[1]
    zr = np.arange(1, N[2] + 1) * pixel_size - c[2]
    Y, X, Z = np.meshgrid(yr, xr, zr, indexing="ij")
    return X, Y, Z, (xr, yr, zr)


def sphere_index(N, pixel_size, radius, n_sphere, n_medium):
    """The refractive-index map of create_sphere.m:39-42 ([y, x, z])."""
    X, Y, Z, _ = reference_grid(N, pixel_size)
    inside = X ** 2 + Y ** 2 + Z ** 2 < radius ** 2
    return np.where(inside, n_sphere, n_medium)
