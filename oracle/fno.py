"""numpy restatement of the learned warm start's network (TEST INFRASTRUCTURE ONLY).

The network has no counterpart in the reference; this module is the CPU
oracle of the project's own HIP forward/backward kernels.  It is pinned by
central finite differences in float64 (tests/test_fno_oracle.py).

Factorised 3-D Fourier neural operator on a cubic grid L^3, activations
[B][C][Z][Y][X] (x fastest):
  lift      h = Wl x + bl                         (Cin -> W channels)
  layer l   h = GELU( sum_a K_a,l h + Wp_l h + b_l ),  a in {x, y, z}
  project   y = Q2 GELU(Q1 h + c1) + c2            (W -> P -> 2)
K_a is a truncated spectral convolution along axis a: the first m Fourier
modes of the real input along that axis (DFT matrix D, [L][2m] = [cos | -sin]),
a per-mode complex channel mixing (Wr_a, Wi_a of shape [m][W][W]) and the
Hermitian inverse DFT restricted to those modes (Dinv, [2m][L], weights
c_k / L with c_0 = 1, c_k = 2).  GELU is the tanh form.

``bf16=True`` rounds both operands of every matrix product to bfloat16
(round to nearest even) as the MFMA kernels do; products accumulate in float64.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


def round_bf16(a: np.ndarray) -> np.ndarray:
    """float32 -> bfloat16 -> float32, round to nearest even."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def dft_matrices(L: int, m: int):
    x = np.arange(L)[:, None]
    k = np.arange(m)[None, :]
    ang = 2 * np.pi * x * k / L
    D = np.concatenate([np.cos(ang), -np.sin(ang)], axis=1)          # [L][2m]
    ck = np.where(np.arange(m) == 0, 1.0, 2.0)[:, None] / L
    Dinv = np.concatenate([ck * np.cos(ang.T), -ck * np.sin(ang.T)], axis=0)  # [2m][L]
    return D, Dinv


def gelu(x):
    s = math.sqrt(2 / math.pi)
    return 0.5 * x * (1 + np.tanh(s * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    s = math.sqrt(2 / math.pi)
    t = np.tanh(s * (x + 0.044715 * x ** 3))
    return 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * s * (1 + 3 * 0.044715 * x * x)


@dataclass
class Config:
    L: int = 16        # grid edge
    cin: int = 4       # input channels
    width: int = 8
    modes: int = 4
    layers: int = 2
    proj: int = 16


def init_params(cfg: Config, seed: int = 0) -> dict:
    rng = np.random.default_rng(seed)
    W, m = cfg.width, cfg.modes
    p = {"Wl": rng.standard_normal((W, cfg.cin)) / math.sqrt(cfg.cin), "bl": np.zeros(W)}
    for l in range(cfg.layers):
        p[f"Wp{l}"] = rng.standard_normal((W, W)) / math.sqrt(W)
        p[f"b{l}"] = np.zeros(W)
        for a in "xyz":
            sc = 1.0 / (W * m)
            p[f"Wr{l}{a}"] = sc * rng.standard_normal((m, W, W))
            p[f"Wi{l}{a}"] = sc * rng.standard_normal((m, W, W))
    p["Q1"] = rng.standard_normal((cfg.proj, W)) / math.sqrt(W)
    p["c1"] = np.zeros(cfg.proj)
    p["Q2"] = rng.standard_normal((2, cfg.proj)) / math.sqrt(cfg.proj)
    p["c2"] = np.zeros(2)
    return p


class Net:
    def __init__(self, cfg: Config, bf16: bool = False):
        self.cfg = cfg
        self.bf16 = bf16
        self.D, self.Dinv = dft_matrices(cfg.L, cfg.modes)

    def mm(self, a, b, subs):
        if self.bf16:
            a, b = round_bf16(a), round_bf16(b)
        return np.einsum(subs, a, b)

    # axis index of [B][C][Z][Y][X]
    AX = {"x": 4, "y": 3, "z": 2}

    def spec_fwd(self, h, Wr, Wi, a):
        ax = self.AX[a]
        hm = np.moveaxis(h, ax, -1)                       # [B][C][..][L]
        F = self.mm(hm, self.D, "bcuvl,lk->bcuvk")        # [B][C][u][v][2m]
        m = self.cfg.modes
        Fr, Fi = F[..., :m], F[..., m:]
        Yr = self.mm(Wr, Fr, "koc,bcuvk->bouvk") - self.mm(Wi, Fi, "koc,bcuvk->bouvk")
        Yi = self.mm(Wi, Fr, "koc,bcuvk->bouvk") + self.mm(Wr, Fi, "koc,bcuvk->bouvk")
        Y = np.concatenate([Yr, Yi], axis=-1)
        out = self.mm(Y, self.Dinv, "bouvk,kl->bouvl")
        return np.moveaxis(out, -1, ax), F

    def spec_bwd(self, dout, F, Wr, Wi, a):
        ax = self.AX[a]
        m = self.cfg.modes
        dm = np.moveaxis(dout, ax, -1)
        dY = self.mm(dm, self.Dinv.T, "bouvl,lk->bouvk")
        dYr, dYi = dY[..., :m], dY[..., m:]
        Fr, Fi = F[..., :m], F[..., m:]
        dWr = self.mm(dYr, Fr, "bouvk,bcuvk->koc") + self.mm(dYi, Fi, "bouvk,bcuvk->koc")
        dWi = -self.mm(dYr, Fi, "bouvk,bcuvk->koc") + self.mm(dYi, Fr, "bouvk,bcuvk->koc")
        dFr = self.mm(Wr, dYr, "koc,bouvk->bcuvk") + self.mm(Wi, dYi, "koc,bouvk->bcuvk")
        dFi = -self.mm(Wi, dYr, "koc,bouvk->bcuvk") + self.mm(Wr, dYi, "koc,bouvk->bcuvk")
        dF = np.concatenate([dFr, dFi], axis=-1)
        dh = self.mm(dF, self.D.T, "bcuvk,kl->bcuvl")
        return np.moveaxis(dh, -1, ax), dWr, dWi

    def forward(self, p, x):
        """x: [B][cin][L][L][L] -> y [B][2][L][L][L]; returns (y, cache)."""
        cfg = self.cfg
        cache = {"x": x}
        h = self.mm(p["Wl"], x, "oc,bczyx->bozyx") + p["bl"][None, :, None, None, None]
        for l in range(cfg.layers):
            cache[f"h{l}"] = h
            pre = self.mm(p[f"Wp{l}"], h, "oc,bczyx->bozyx")
            for a in "xyz":
                s, F = self.spec_fwd(h, p[f"Wr{l}{a}"], p[f"Wi{l}{a}"], a)
                cache[f"F{l}{a}"] = F
                pre = pre + s
            pre = pre + p[f"b{l}"][None, :, None, None, None]
            cache[f"pre{l}"] = pre
            h = gelu(pre)
        cache["hL"] = h
        qp = self.mm(p["Q1"], h, "oc,bczyx->bozyx") + p["c1"][None, :, None, None, None]
        cache["qp"] = qp
        q = gelu(qp)
        cache["q"] = q
        y = self.mm(p["Q2"], q, "oc,bczyx->bozyx") + p["c2"][None, :, None, None, None]
        return y, cache

    def backward(self, p, cache, dy):
        cfg = self.cfg
        g = {}
        sumv = lambda t: t.sum(axis=(0, 2, 3, 4))
        g["Q2"] = self.mm(dy, cache["q"], "bozyx,bczyx->oc")
        g["c2"] = sumv(dy)
        dq = self.mm(p["Q2"], dy, "oc,bozyx->bczyx")
        dqp = dq * gelu_grad(cache["qp"])
        g["Q1"] = self.mm(dqp, cache["hL"], "bozyx,bczyx->oc")
        g["c1"] = sumv(dqp)
        dh = self.mm(p["Q1"], dqp, "oc,bozyx->bczyx")
        for l in reversed(range(cfg.layers)):
            dpre = dh * gelu_grad(cache[f"pre{l}"])
            g[f"b{l}"] = sumv(dpre)
            h = cache[f"h{l}"]
            g[f"Wp{l}"] = self.mm(dpre, h, "bozyx,bczyx->oc")
            dh = self.mm(p[f"Wp{l}"], dpre, "oc,bozyx->bczyx")
            for a in "xyz":
                dha, dWr, dWi = self.spec_bwd(dpre, cache[f"F{l}{a}"], p[f"Wr{l}{a}"], p[f"Wi{l}{a}"], a)
                g[f"Wr{l}{a}"] = dWr
                g[f"Wi{l}{a}"] = dWi
                dh = dh + dha
        g["Wl"] = self.mm(dh, cache["x"], "bozyx,bczyx->oc")
        g["bl"] = sumv(dh)
        return g

    def loss(self, p, x, target):
        """0.5 * mean((y - target)^2) and its gradient."""
        y, cache = self.forward(p, x)
        n = y.size
        e = y - target
        L = 0.5 * float(np.sum(e * e)) / n
        g = self.backward(p, cache, e / n)
        return L, g, y
