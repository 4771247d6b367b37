"""Learned warm start for plane-wave problems (no counterpart in the reference).

The network predicts the envelope A of the field relative to the closed-form
background solution u(z) of the source plane (``prior.green_1d``):
    E0 = u(z) (1 + A)  inside the ROI,   u(z) on the source footprint outside.
Input features on a grid pooled by ``factor`` from the ROI: the permittivity
contrast n^2 - n_bg^2 and the three normalised coordinates.  The network runs
on the project's HIP kernels (``fno.FNO``), the field is composed on the
solver grid by ``wfno_compose_warm`` and handed to the MBS solver as a warm
start (``wsim_solver_set_warm_start``); the plain MBS iteration then runs to
the same absolute residual as a zero start, so zero-start MBS is both the
fallback and the judge of the warm start's error.

Training data are generated by the solver itself (zero-start solves of
seeded sphere media, ``workloads.spheres_3d``).
"""
from __future__ import annotations

import concurrent.futures as cf
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import kernels as K
from .fno import FNO, Config
from .prior import green_1d


@dataclass
class Sample:
    x: np.ndarray        # [cin][L^3] features
    t: np.ndarray        # [2][L^3] target envelope
    iterations: int
    seconds: float


class WarmStart:
    def __init__(self, cfg: Config | None = None, factor: int = 2, n_bg: float = 1.33, params=None,
                 seed: int = 0):
        self.cfg = cfg or Config()
        self.factor = factor
        self.n_bg = n_bg
        self.net = FNO(self.cfg, params, seed=seed)

    # -- geometry ------------------------------------------------------------------
    @staticmethod
    def roi_slices(sim):
        r = sim.roi
        return tuple(slice(int(r[0, d]) - 1, int(r[1, d])) for d in range(3))

    def prior_u(self, sim) -> np.ndarray:
        """u(z) on the grid's z axis for the source plane at the ROI's first z."""
        zs = int(sim.roi[0, 2]) - 1
        k = sim.k0 * self.n_bg
        return green_1d(k, sim.grid.dx, (np.arange(sim.N[2]) - zs) * sim.grid.dx).astype(np.complex64)

    def _pool(self, a):
        f = self.factor
        Z, Y, X = a.shape
        return a.reshape(Z // f, f, Y // f, f, X // f, f).mean(axis=(1, 3, 5))

    def features(self, nmap_roi) -> np.ndarray:
        """nmap_roi: [y][x][z] refractive index of the ROI -> [cin][L^3]."""
        d = (np.real(nmap_roi).astype(np.float32) ** 2 - np.float32(self.n_bg ** 2)).transpose(2, 0, 1)
        d = self._pool(d)
        L = d.shape[0]
        assert d.shape == (L, L, L) and L == self.cfg.L, (d.shape, self.cfg.L)
        c = (np.arange(L, dtype=np.float32) + 0.5) / L - 0.5
        cz, cy, cx = np.meshgrid(c, c, c, indexing="ij")
        return np.stack([d, cz, cy, cx]).reshape(4, -1).astype(np.float32)

    def target(self, sim, E_roi) -> np.ndarray:
        """Envelope of the solved field E_roi ([y][x][z] complex, ROI) relative to u."""
        u = self.prior_u(sim)[self.roi_slices(sim)[2]]
        A = (E_roi / u[None, None, :] - 1).transpose(2, 0, 1)
        A = self._pool(A)
        return np.stack([A.real, A.imag]).reshape(2, -1).astype(np.float32)

    # -- inference -----------------------------------------------------------------
    def predict_E0(self, sim, x: np.ndarray | None = None, nmap_roi=None) -> torch.Tensor:
        """Warm-start field on the solver grid (device, solver layout)."""
        if x is None:
            x = self.features(nmap_roi)
        xt = torch.from_numpy(x).to(self.net.dev).reshape(1, self.cfg.cin, -1)
        y, _ = self.net.forward(xt, keep=False)
        return self.compose(sim, y)

    def compose(self, sim, y: torch.Tensor) -> torch.Tensor:
        import ctypes as C
        u = torch.from_numpy(self.prior_u(sim).view(np.float32).copy()).to(self.net.dev)
        E0 = torch.empty(2 * int(np.prod(sim.N[:3])), dtype=torch.float32, device=self.net.dev)
        sl = self.roi_slices(sim)
        grid = (C.c_int * 3)(*sim.N[:3])
        lo = (C.c_int * 3)(*[s.start for s in sl])
        n = (C.c_int * 3)(*[s.stop - s.start for s in sl])
        foot = (C.c_int * 4)(sl[0].start, sl[0].stop, sl[1].start, sl[1].stop)
        K.check(K.lib().wfno_compose_warm(K.ptr(E0), K.ptr(y), K.ptr(u), grid, lo, n, self.factor, foot,
                                          K.stream_ptr()))
        return E0

    # -- training ------------------------------------------------------------------
    def train(self, X: torch.Tensor, T: torch.Tensor, epochs: int = 10, batch: int = 4, lr: float = 2e-3,
              log=None, seed: int = 0):
        """X: [n][cin][L^3], T: [n][2][L^3] device tensors.  Cosine-decayed Adam."""
        n = X.shape[0]
        rng = np.random.default_rng(seed)
        steps = epochs * max(1, n // batch)
        step = 0
        hist = []
        for ep in range(epochs):
            perm = rng.permutation(n)
            tot, cnt = 0.0, 0
            for i in range(0, n - batch + 1, batch):
                idx = torch.from_numpy(perm[i:i + batch]).to(X.device)
                xb = X.index_select(0, idx).contiguous()
                tb = T.index_select(0, idx).contiguous()
                loss, _ = self.net.loss_and_grad(xb, tb)
                a = lr * 0.5 * (1 + np.cos(np.pi * step / steps))
                self.net.adam_step(lr=float(a))
                tot += loss
                cnt += 1
                step += 1
            hist.append(tot / max(cnt, 1))
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.5f}")
        return hist

    def save(self, path: str):
        np.savez(path, **{k: v.astype(np.float32) for k, v in self.net.params().items()},
                 _cfg=np.array([self.cfg.L, self.cfg.cin, self.cfg.width, self.cfg.modes, self.cfg.layers,
                                self.cfg.proj, self.factor], np.int64),
                 _nbg=np.array([self.n_bg]))

    @classmethod
    def load(cls, path: str) -> "WarmStart":
        d = np.load(path, allow_pickle=False)
        c = [int(v) for v in d["_cfg"]]
        cfg = Config(L=c[0], cin=c[1], width=c[2], modes=c[3], layers=c[4], proj=c[5])
        params = {k: d[k] for k in d.files if not k.startswith("_")}
        return cls(cfg, factor=c[6], n_bg=float(d["_nbg"][0]), params=params)


# ----------------------------------------------------------------------------------
# data generation and A/B evaluation
# ----------------------------------------------------------------------------------
def _build(seed, n, boundary):
    from .. import WaveSim
    from ..workloads import spheres_3d
    nmap, opt, src = spheres_3d(n, boundary, seed=seed)
    sim = WaveSim(nmap, opt)
    return seed, nmap, sim, src


def iter_media(seeds, n=128, boundary=32, workers=8):
    """Yields (seed, nmap, sim, source) with the host set-up of later seeds
    running in a thread pool while the caller uses the GPU."""
    with cf.ThreadPoolExecutor(workers) as ex:
        futs = [ex.submit(_build, s, n, boundary) for s in seeds[:workers]]
        nxt = workers
        for i in range(len(seeds)):
            r = futs[i].result()
            futs[i] = None
            if nxt < len(seeds):
                futs.append(ex.submit(_build, seeds[nxt], n, boundary))
                nxt += 1
            yield r


def generate(ws: WarmStart, seeds, n=128, boundary=32, log=None):
    """Zero-start solves of seeded sphere media -> features and targets."""
    X, T, info = [], [], []
    holder = None
    for seed, nmap, sim, src in iter_media(list(seeds), n, boundary):
        if holder is None:
            sv = sim.solver(1)
        else:
            sv = sim.adopt_solver(holder)
        holder = sim
        p = sim._prepare(src)
        sv.set_source(0, p.positions, p.values)
        t0 = time.perf_counter()
        sv.solve(sim.max_iterations(), sim.energy_threshold)
        dt = time.perf_counter() - t0
        r = sv.result(0)
        with np.errstate(all="ignore"):
            E = sim.crop_field(sv.field(0) / sim.gamma[..., None])[..., 0]
        X.append(ws.features(nmap))
        T.append(ws.target(sim, E))
        info.append((seed, r.iterations, dt))
        if log and len(info) % 50 == 0:
            log(f"generated {len(info)} samples (last: {r.iterations} iterations, {dt * 1e3:.1f} ms)")
    return np.stack(X), np.stack(T), info, holder


def evaluate(ws: WarmStart, seeds, holder=None, n=128, boundary=32, tight=1e-10):
    """Per seed: zero start, prior-only warm start and learned warm start, each
    run to the zero start's absolute residual (same stop denominator), timed
    from inputs resident on the GPU; errors against a tight zero-start solve."""
    rows = []
    for seed, nmap, sim, src in iter_media(list(seeds), n, boundary, workers=4):
        sv = sim.solver(1) if holder is None else sim.adopt_solver(holder)
        holder = sim
        p = sim._prepare(src)
        sv.set_source(0, p.positions, p.values)
        mi = sim.max_iterations()
        g4 = sim.gamma[..., None]

        def field():
            with np.errstate(all="ignore"):
                return sim.crop_field(sv.field(0) / g4)

        sv.solve(mi, tight)
        Eref = field()
        sv.solve(mi, sim.energy_threshold)           # warm-up of this medium
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sv.solve(mi, sim.energy_threshold)
        t_zero = time.perf_counter() - t0
        rz = sv.result(0)
        Ez = field()
        d1 = float(rz.diff_energy[0])
        err = lambda E: float(np.linalg.norm(E - Eref) / np.linalg.norm(Eref))
        out = dict(seed=int(seed), zero_its=rz.iterations, zero_s=t_zero, zero_err=err(Ez))
        x = ws.features(nmap)
        for name in ("prior", "learned"):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if name == "learned":
                E0 = ws.predict_E0(sim, x=x)
            else:
                zero = torch.zeros(2 * ws.cfg.L ** 3, dtype=torch.float32, device=ws.net.dev)
                E0 = ws.compose(sim, zero)
            sv.set_warm_start(0, device_ptr=K.ptr(E0))
            sv.set_reference_energy(0, d1)
            t_pred = time.perf_counter() - t0
            sv.solve(mi, sim.energy_threshold)
            torch.cuda.synchronize()
            t_tot = time.perf_counter() - t0
            r = sv.result(0)
            Ew = field()
            out[f"{name}_its"] = r.iterations
            out[f"{name}_s"] = t_tot
            out[f"{name}_predict_s"] = t_pred
            out[f"{name}_err"] = err(Ew)
            sv.clear_warm_start()
            sv.set_reference_energy(0, 0.0)
        rows.append(out)
    return rows, holder
