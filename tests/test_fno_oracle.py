"""Pins for the network oracle: the truncated DFT matrices against numpy.fft,
the spectral convolution against an FFT formulation, and every gradient of
the explicit backward pass against central finite differences (float64)."""
import numpy as np
import pytest

from oracle import fno


def test_dft_matrices_match_numpy_fft():
    L, m = 16, 5
    D, Dinv = fno.dft_matrices(L, m)
    rng = np.random.default_rng(0)
    h = rng.standard_normal(L)
    F = h @ D
    ref = np.fft.rfft(h)[:m]
    assert np.allclose(F[:m] + 1j * F[m:], ref)
    # inverse: irfft of the first m modes (others zero)
    Y = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    Y[0] = Y[0].real
    full = np.zeros(L // 2 + 1, complex)
    full[:m] = Y
    assert np.allclose(np.concatenate([Y.real, Y.imag]) @ Dinv, np.fft.irfft(full, L))


def test_spectral_axis_equals_fft_filter():
    cfg = fno.Config(L=8, cin=2, width=3, modes=3, layers=1, proj=4)
    net = fno.Net(cfg)
    rng = np.random.default_rng(1)
    h = rng.standard_normal((2, 3, 8, 8, 8))
    Wr = rng.standard_normal((3, 3, 3))
    Wi = rng.standard_normal((3, 3, 3))
    out, _ = net.spec_fwd(h, Wr, Wi, "y")
    H = np.fft.rfft(h, axis=3)[:, :, :, :3, :]            # modes along y
    Wc = Wr + 1j * Wi                                     # [k][o][c]
    Yk = np.einsum("koc,bczkx->bozkx", Wc, H)
    full = np.zeros((2, 3, 8, 5, 8), complex)
    full[:, :, :, :3, :] = Yk
    ref = np.fft.irfft(full, 8, axis=3)
    assert np.allclose(out, ref)


def test_gradients_finite_differences():
    cfg = fno.Config(L=8, cin=3, width=4, modes=3, layers=2, proj=5)
    net = fno.Net(cfg)
    p = fno.init_params(cfg, seed=2)
    rng = np.random.default_rng(3)
    for k in p:
        p[k] = p[k] + 0.1 * rng.standard_normal(p[k].shape)
    x = rng.standard_normal((2, 3, 8, 8, 8))
    t = rng.standard_normal((2, 2, 8, 8, 8))
    L0, g, _ = net.loss(p, x, t)
    eps = 1e-6
    for name in p:
        flat = p[name].reshape(-1)
        idx = rng.choice(flat.size, size=min(4, flat.size), replace=False)
        for i in idx:
            old = flat[i]
            flat[i] = old + eps
            Lp, _, _ = net.loss(p, x, t)
            flat[i] = old - eps
            Lm, _, _ = net.loss(p, x, t)
            flat[i] = old
            fd = (Lp - Lm) / (2 * eps)
            an = g[name].reshape(-1)[i]
            assert abs(fd - an) <= 1e-6 + 1e-5 * abs(fd), (name, i, fd, an)


def test_bf16_rounding_round_to_nearest_even():
    # bfloat16 keeps 8 significant bits: the spacing above 1.0 is 2^-7
    vals = np.array([1.0, 1 + 2 ** -8, 1 + 2 ** -9, 1 + 3 * 2 ** -9, 1 + 3 * 2 ** -8, -3.0000002],
                    np.float32)
    r = fno.round_bf16(vals)
    # 1 + 2^-8: tie -> even (1.0); 1 + 2^-9: below half -> 1.0;
    # 1 + 3*2^-9: above half -> 1 + 2^-7; 1 + 3*2^-8: tie -> even (1 + 2^-6)
    assert list(r[:5]) == [1.0, 1.0, 1.0, 1 + 2 ** -7, 1 + 2 ** -6]
    assert r[5] == -3.0


def test_network_dft_matrices_interleave_real_and_imaginary_parts():
    """The GPU network's truncated DFT keeps each mode's real and imaginary
    part adjacent (columns 2k, 2k+1), so that one GEMM with the 2W x 2W real
    form of the complex weight mixes a mode; checked against numpy.fft."""
    from wavesim_amd.learned.fno import dft_matrices
    L, m = 16, 5
    D, Dinv = dft_matrices(L, m)
    assert D.shape == (L, 2 * m) and Dinv.shape == (2 * m, L)
    rng = np.random.default_rng(2)
    h = rng.standard_normal(L)
    F = h @ D
    assert np.allclose(F[0::2] + 1j * F[1::2], np.fft.rfft(h)[:m])
    Y = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    Y[0] = Y[0].real
    full = np.zeros(L // 2 + 1, complex)
    full[:m] = Y
    inter = np.stack([Y.real, Y.imag], axis=1).ravel()
    assert np.allclose(inter @ Dinv, np.fft.irfft(full, L))
