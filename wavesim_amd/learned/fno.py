"""Factorised 3-D Fourier neural operator on the project's HIP kernels.

Architecture (cubic grid L^3, activations [B][C][Z][Y][X], x fastest):
  lift      h = Wl x + bl                          (cin -> width)
  layer l   h = GELU( sum_a K_a,l h + Wp_l h + b_l ),  a in {x, y, z}
  project   y = Q2 GELU(Q1 h + c1) + c2            (width -> proj -> 2)
K_a: first m Fourier modes along axis a (truncated DFT as a GEMM with the
[L][2m] matrix D whose columns 2k, 2k+1 are cos and -sin of mode k),
per-mode complex channel mixing (Wr, Wi of shape [m][W][W]; one real GEMM per
mode with the 2W x 2W weight [[Wr, -Wi], [Wi, Wr]] on the mode's real and
imaginary parts, which the interleaved columns of D keep adjacent), Hermitian
inverse DFT over those modes (GEMM with Dinv [2m][L], rows interleaved alike).
This is the factorised ("F-FNO") spectral layer: 3 * 2 * m * W^2 weights per
layer instead of 2 * m^3 * W^2 for a dense 3-D mode block.

Every matrix product (forward and backward) is one launch of wfno_gemm (BF16
MFMA, FP32 accumulate); bias, GELU, loss, Adam and the assembly of the
2W x 2W mixing weights are the project's elementwise kernels.  torch tensors
only own device memory.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import kernels as K


@dataclass
class Config:
    L: int = 64
    cin: int = 4
    width: int = 32
    modes: int = 16
    layers: int = 4
    proj: int = 64


def dft_matrices(L: int, m: int):
    """D [L][2m]: column 2k = cos, 2k+1 = -sin of mode k (real and imaginary
    part of the truncated DFT); Dinv [2m][L]: the Hermitian inverse over
    those modes, rows in the same interleaved order."""
    x = np.arange(L)[:, None]
    k = np.arange(m)[None, :]
    ang = 2 * np.pi * x * k / L
    D = np.stack([np.cos(ang), -np.sin(ang)], axis=2).reshape(L, 2 * m)
    ck = np.where(np.arange(m) == 0, 1.0, 2.0)[:, None] / L
    Dinv = np.stack([ck * np.cos(ang.T), -ck * np.sin(ang.T)], axis=1).reshape(2 * m, L)
    return D, Dinv


def param_shapes(cfg: Config):
    W, m = cfg.width, cfg.modes
    s = [("Wl", (W, cfg.cin)), ("bl", (W,))]
    for l in range(cfg.layers):
        s += [(f"Wp{l}", (W, W)), (f"b{l}", (W,))]
        for a in "xyz":
            s += [(f"Wr{l}{a}", (m, W, W)), (f"Wi{l}{a}", (m, W, W))]
    s += [("Q1", (cfg.proj, W)), ("c1", (cfg.proj,)), ("Q2", (2, cfg.proj)), ("c2", (2,))]
    return s


def init_params(cfg: Config, seed: int = 0) -> dict:
    rng = np.random.default_rng(seed)
    W, m = cfg.width, cfg.modes
    p = {}
    for name, shape in param_shapes(cfg):
        if name.startswith("b") or name.startswith("c"):
            p[name] = np.zeros(shape)
        elif name.startswith("Wr") or name.startswith("Wi"):
            p[name] = rng.standard_normal(shape) / (W * m)
        else:
            p[name] = rng.standard_normal(shape) / math.sqrt(shape[1])
    return p


def forward_gemm_roles(cfg: Config):
    """What each wfno_gemm launch of FNO.forward computes, in launch order
    (for reading kernel traces of a forward)."""
    roles = ["lift"]
    for _ in range(cfg.layers):
        roles.append("pointwise Wp")
        for a in "xyz":
            roles += [f"truncated DFT {a}", "complex mixing", f"inverse truncated DFT {a}"]
    return roles + ["projection Q1", "projection Q2"]


class FNO:
    """Parameters live in one flat device buffer (one Adam launch)."""

    def __init__(self, cfg: Config, params: dict | None = None, device="cuda", seed: int = 0):
        self.cfg = cfg
        self.dev = torch.device(device)
        self.shapes = param_shapes(cfg)
        self.offsets = {}
        off = 0
        for name, shape in self.shapes:
            self.offsets[name] = (off, shape)
            off += int(np.prod(shape))
        self.nparam = off
        self.flat = torch.zeros(off, dtype=torch.float32, device=self.dev)
        self.grad = torch.zeros_like(self.flat)
        self.m = torch.zeros_like(self.flat)
        self.v = torch.zeros_like(self.flat)
        self.t = 0
        self.set_params(params if params is not None else init_params(cfg, seed))
        D, Dinv = dft_matrices(cfg.L, cfg.modes)
        self.D = torch.tensor(D, dtype=torch.float32, device=self.dev).contiguous()
        self.Dinv = torch.tensor(Dinv, dtype=torch.float32, device=self.dev).contiguous()
        self._bufs = {}

    # -- parameters ----------------------------------------------------------------
    def set_params(self, p: dict):
        host = np.zeros(self.nparam, np.float32)
        for name, (off, shape) in self.offsets.items():
            host[off:off + int(np.prod(shape))] = np.asarray(p[name], np.float32).ravel()
        self.flat.copy_(torch.from_numpy(host))

    def params(self) -> dict:
        host = self.flat.cpu().numpy()
        return {n: host[o:o + int(np.prod(s))].reshape(s).astype(np.float64) for n, (o, s) in self.offsets.items()}

    def grads(self) -> dict:
        host = self.grad.cpu().numpy()
        return {n: host[o:o + int(np.prod(s))].reshape(s).astype(np.float64) for n, (o, s) in self.offsets.items()}

    def P(self, name):
        return (self.flat, self.offsets[name][0])

    def G(self, name):
        return (self.grad, self.offsets[name][0])

    def buf(self, key, n):
        t = self._bufs.get(key)
        if t is None or t.numel() < n:
            t = torch.empty(n, dtype=torch.float32, device=self.dev)
            self._bufs[key] = t
        return t

    # -- building blocks -------------------------------------------------------------
    def pointwise(self, Wt, X, Y, cout, cin, B, beta=0.0):
        """Y[b] (cout x V) = W (cout x cin) X[b] (cin x V) + beta Y[b]."""
        V = self.cfg.L ** 3
        K.gemm(cout, V, cin, Wt, (cin, 1, 0), X, (V, 1, cin * V), Y, (V, 1, cout * V), beta=beta, nb=(B,))

    def pointwise_dW(self, dY, X, dW, cout, cin, B, accumulate=False):
        """dW (cout x cin) (+)= sum_b dY[b] X[b]^T, split over V in chunks."""
        V = self.cfg.L ** 3
        nch = 1  # split V into nch equal chunks (K of each GEMM >= 1024)
        for d in range(min(64, V // 1024), 0, -1):
            if V % d == 0:
                nch = d
                break
        kc = V // nch
        tmp = self.buf("dw_tmp", B * nch * cout * cin)
        # C[b][chunk] = dY[b][:, chunk] X[b][:, chunk]^T
        K.gemm(cout, cin, kc, dY, (V, 1, cout * V, kc), X, (1, V, cin * V, kc),
               tmp, (cin, 1, nch * cout * cin, cout * cin), nb=(B, nch))
        K.check(K.lib().wfno_reduce_chunks(K.ptr(tmp), K._addr(dW), B * nch, cout * cin, int(accumulate),
                                           K.stream_ptr()))

    def bias_act(self, pre, out, bias, C, B, act):
        V = self.cfg.L ** 3
        K.check(K.lib().wfno_bias_act(K._addr(pre), K._addr(out), K._addr(bias), C, V, B, int(act), K.stream_ptr()))

    def bias_grad(self, d, dbias, C, B):
        V = self.cfg.L ** 3
        K.check(K.lib().wfno_bias_grad(K._addr(d), K._addr(dbias), C, V, B, 0, K.stream_ptr()))

    def gelu_bwd(self, dout, pre, dpre, n):
        K.check(K.lib().wfno_gelu_bwd(K._addr(dout), K._addr(pre), K._addr(dpre), n, K.stream_ptr()))

    # transforms along an axis: "fwd" maps h-layout [B][C][L^3] to the mode
    # layout [B][m][2][C][L^2] (real, imaginary part of each mode adjacent);
    # "inv" maps back (accumulating)
    def to_modes(self, a, Mat, s_rk, s_x, src, dst, C, B):
        L, m = self.cfg.L, self.cfg.modes
        V, L2, r = L ** 3, L * L, 2 * self.cfg.modes
        if a == "x":
            K.gemm(r, C * L2, L, Mat, (s_rk, s_x), src, (1, L, C * V), dst, (C * L2, 1, r * C * L2), nb=(B,))
        elif a == "y":
            K.gemm(r, L, L, Mat, (s_rk, s_x, 0, 0, 0), src, (L, 1, C * V, V, L2),
                   dst, (C * L2, 1, r * C * L2, L2, L), nb=(B, C, L))
        else:
            K.gemm(r, L2, L, Mat, (s_rk, s_x, 0, 0), src, (L2, 1, C * V, V),
                   dst, (C * L2, 1, r * C * L2, L2), nb=(B, C))

    def from_modes(self, a, Mat, s_rk, s_x, src, dst, C, B, beta=1.0):
        L, m = self.cfg.L, self.cfg.modes
        V, L2, r = L ** 3, L * L, 2 * self.cfg.modes
        if a == "x":
            K.gemm(C * L2, L, r, src, (1, C * L2, r * C * L2), Mat, (s_rk, s_x, 0),
                   dst, (L, 1, C * V), beta=beta, nb=(B,))
        elif a == "y":
            K.gemm(L, L, r, Mat, (s_x, s_rk, 0, 0, 0), src, (C * L2, 1, r * C * L2, L2, L),
                   dst, (L, 1, C * V, V, L2), beta=beta, nb=(B, C, L))
        else:
            K.gemm(L, L2, r, Mat, (s_x, s_rk, 0, 0), src, (C * L2, 1, r * C * L2, L2),
                   dst, (L2, 1, C * V, V), beta=beta, nb=(B, C))

    def complex_weights(self):
        """Wc [layers][3][m][2W][2W] = [[Wr, -Wi], [Wi, Wr]] per layer, axis
        and mode, from the current parameters (one launch)."""
        cfg = self.cfg
        W, m = cfg.width, cfg.modes
        Wc = self.buf("Wc", cfg.layers * 3 * m * 4 * W * W)
        o0 = self.offsets["Wr0x"][0]
        pair = self.offsets["Wr0y"][0] - o0
        group = self.offsets["Wr1x"][0] - o0 if cfg.layers > 1 else 3 * pair
        for l in range(cfg.layers):
            for j, a in enumerate("xyz"):
                assert self.offsets[f"Wr{l}{a}"][0] == o0 + l * group + j * pair
                assert self.offsets[f"Wi{l}{a}"][0] == o0 + l * group + j * pair + m * W * W
        K.check(K.lib().wfno_complex_weights(K._addr((self.flat, o0)), K.ptr(Wc), pair, 3, group,
                                             3 * cfg.layers, m, W, K.stream_ptr()))
        return Wc

    def mix(self, Wc, F, Y, B, transpose=False):
        """Per-mode complex channel mixing in the mode layout [B][m][2][W][L^2]:
        [Yr; Yi]_k = Wc_k [Fr; Fi]_k with Wc_k = [[Wr, -Wi], [Wi, Wr]], i.e.
        Y = (Wr + i Wi) F, one GEMM (2W x L^2 x 2W per mode); transpose=True
        applies Wc_k^T, the adjoint dF = Wr^T dY_r + Wi^T dY_i,
        Wr^T dY_i - Wi^T dY_r."""
        W, m, L2 = self.cfg.width, self.cfg.modes, self.cfg.L ** 2
        sW = (1, 2 * W, 0, 4 * W * W) if transpose else (2 * W, 1, 0, 4 * W * W)
        sF = (L2, 1, 2 * m * W * L2, 2 * W * L2)
        K.gemm(2 * W, L2, 2 * W, Wc, sW, F, sF, Y, sF, nb=(B, m))

    def mix_dW(self, dY, F, dWr, dWi, B):
        """dWr = sum dYr Fr^T + dYi Fi^T,  dWi = sum dYi Fr^T - dYr Fi^T."""
        W, m, L2 = self.cfg.width, self.cfg.modes, self.cfg.L ** 2
        im = W * L2                       # imaginary part of a mode right after its real part
        sA = (L2, 1, 0, 2 * W * L2)       # dY_k[b] as (o, line)
        sB = (1, L2, 0, 2 * W * L2)       # F_k[b]^T as (line, c)
        sC = (W, 1, 0, W * W)
        for b in range(B):
            o = b * 2 * m * W * L2
            first = b == 0
            K.gemm(W, W, L2, (dY, o), sA, (F, o), sB, dWr, sC, 1.0, 0.0 if first else 1.0, nb=(1, m))
            K.gemm(W, W, L2, (dY, o + im), sA, (F, o + im), sB, dWr, sC, 1.0, 1.0, nb=(1, m))
            K.gemm(W, W, L2, (dY, o + im), sA, (F, o), sB, dWi, sC, 1.0, 0.0 if first else 1.0, nb=(1, m))
            K.gemm(W, W, L2, (dY, o), sA, (F, o + im), sB, dWi, sC, -1.0, 1.0, nb=(1, m))

    # -- network ---------------------------------------------------------------------
    def forward(self, x: torch.Tensor, keep: bool = True):
        """x: [B][cin][L^3] float32 -> y [B][2][L^3]."""
        cfg = self.cfg
        B = x.shape[0]
        W, V, L2, r = cfg.width, cfg.L ** 3, cfg.L ** 2, 2 * cfg.modes
        D, Dinv = self.D, self.Dinv
        s_D = (1, r)        # D[x][rk] as Mat(rk, x)
        s_Di = (cfg.L, 1)   # Dinv[rk][x]
        cache = {"x": x, "B": B}
        Wc = self.complex_weights()
        nc = 4 * W * W * cfg.modes
        h = torch.empty(B * W * V, device=self.dev)
        self.pointwise(self.P("Wl"), x, h, W, cfg.cin, B)
        self.bias_act(h, h, self.P("bl"), W, B, act=False)
        Y = self.buf("Y", B * r * W * L2)
        for l in range(cfg.layers):
            cache[f"h{l}"] = h
            pre = torch.empty(B * W * V, device=self.dev)
            self.pointwise(self.P(f"Wp{l}"), h, pre, W, W, B)
            for j, a in enumerate("xyz"):
                F = torch.empty(B * r * W * L2, device=self.dev) if keep else self.buf("F", B * r * W * L2)
                self.to_modes(a, D, *s_D, h, F, W, B)
                self.mix((Wc, (3 * l + j) * nc), F, Y, B)
                self.from_modes(a, Dinv, *s_Di, Y, pre, W, B, beta=1.0)
                if keep:
                    cache[f"F{l}{a}"] = F
            hn = torch.empty(B * W * V, device=self.dev)
            self.bias_act(pre, hn, self.P(f"b{l}"), W, B, act=True)
            cache[f"pre{l}"] = pre
            h = hn
        cache["hL"] = h
        qp = torch.empty(B * cfg.proj * V, device=self.dev)
        self.pointwise(self.P("Q1"), h, qp, cfg.proj, W, B)
        q = torch.empty_like(qp)
        self.bias_act(qp, q, self.P("c1"), cfg.proj, B, act=True)
        cache["qp"], cache["q"] = qp, q
        y = torch.empty(B * 2 * V, device=self.dev)
        self.pointwise(self.P("Q2"), q, y, 2, cfg.proj, B)
        self.bias_act(y, y, self.P("c2"), 2, B, act=False)
        return y.view(B, 2, V), cache

    def backward(self, cache, dy: torch.Tensor):
        cfg = self.cfg
        B = cache["B"]
        W, V, L2, r = cfg.width, cfg.L ** 3, cfg.L ** 2, 2 * cfg.modes
        s_DT = (cfg.L, 1)     # Dinv[rk][x] as Mat(rk, x)  (transpose of the inverse)
        s_D = (1, r)          # D[x][rk]                    (transpose of the forward)
        dy = dy.reshape(-1)
        self.pointwise_dW(dy, cache["q"], self.G("Q2"), 2, cfg.proj, B)
        self.bias_grad(dy, self.G("c2"), 2, B)
        dq = self.buf("dq", B * cfg.proj * V)
        K.gemm(cfg.proj, V, 2, self.P("Q2"), (1, cfg.proj, 0), dy, (V, 1, 2 * V), dq, (V, 1, cfg.proj * V),
               nb=(B,))
        self.gelu_bwd(dq, cache["qp"], dq, B * cfg.proj * V)
        self.pointwise_dW(dq, cache["hL"], self.G("Q1"), cfg.proj, W, B)
        self.bias_grad(dq, self.G("c1"), cfg.proj, B)
        dh = self.buf("dh", B * W * V)
        K.gemm(W, V, cfg.proj, self.P("Q1"), (1, W, 0), dq, (V, 1, cfg.proj * V), dh, (V, 1, W * V), nb=(B,))
        dpre = self.buf("dpre", B * W * V)
        dY = self.buf("dY", B * r * W * L2)
        dF = self.buf("dF", B * r * W * L2)
        Wc = self.complex_weights()
        nc = 4 * W * W * cfg.modes
        for l in reversed(range(cfg.layers)):
            self.gelu_bwd(dh, cache[f"pre{l}"], dpre, B * W * V)
            self.bias_grad(dpre, self.G(f"b{l}"), W, B)
            h = cache[f"h{l}"]
            self.pointwise_dW(dpre, h, self.G(f"Wp{l}"), W, W, B)
            # dh = Wp^T dpre + sum_a spectral adjoints
            K.gemm(W, V, W, self.P(f"Wp{l}"), (1, W, 0), dpre, (V, 1, W * V), dh, (V, 1, W * V), nb=(B,))
            for j, a in enumerate("xyz"):
                F = cache[f"F{l}{a}"]
                self.to_modes(a, self.Dinv, *s_DT, dpre, dY, W, B)
                self.mix_dW(dY, F, self.G(f"Wr{l}{a}"), self.G(f"Wi{l}{a}"), B)
                self.mix((Wc, (3 * l + j) * nc), dY, dF, B, transpose=True)
                self.from_modes(a, self.D, *s_D, dF, dh, W, B, beta=1.0)
        self.pointwise_dW(dh, cache["x"], self.G("Wl"), W, cfg.cin, B)
        self.bias_grad(dh, self.G("bl"), W, B)

    def loss_and_grad(self, x, target):
        """0.5 mean((y - target)^2); leaves the gradient in self.grad."""
        y, cache = self.forward(x, keep=True)
        n = y.numel()
        nblk = 256
        part = torch.empty(nblk, dtype=torch.float64, device=self.dev)
        dy = self.buf("dy", n)
        K.check(K.lib().wfno_mse(K.ptr(y), K.ptr(target), K.ptr(dy), K.ptr(part), nblk, n, 1.0 / n,
                                 K.stream_ptr()))
        self.backward(cache, dy[:n])
        return 0.5 * float(part.sum().item()) / n, y

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.t += 1
        K.check(K.lib().wfno_adam(K.ptr(self.flat), K.ptr(self.grad), K.ptr(self.m), K.ptr(self.v),
                                  self.nparam, lr, b1, b2, eps, self.t, K.stream_ptr()))
