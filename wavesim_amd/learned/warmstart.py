"""Learned warm start for plane-wave problems (no counterpart in the reference).

Prior: the straight-ray (WKB) field of the source plane, computed at full
resolution on the GPU (``wfno_wkb_prior``): the closed-form background field
u(z) of the plane (``prior.green_1d``, the reference utility
homogeneous_medium_analytic_solution.m:26-35) times the phase
exp(i k0 int (n - n_bg) dz) accumulated along each ray.  The network predicts
the envelope A of the field relative to that prior,
    E0 = P (1 + A)  inside the ROI,   u(z) on the source footprint outside,
from features on a grid pooled by ``factor`` (``wfno_features``): contrast
n^2 - n_bg^2, cos/sin of the ray phase and the normalised coordinates.  The
network runs on the project's HIP kernels (``fno.FNO``); E0 is handed to the
MBS solver as a warm start (``wsim_solver_set_warm_start``) and the plain MBS
iteration runs to the same absolute residual as a zero start, so zero-start
MBS is both the fallback and the judge of the warm start's error.

Training data are generated by the solver itself (zero-start solves of
seeded sphere media, ``workloads.spheres_3d``).
"""
from __future__ import annotations

import concurrent.futures as cf
import ctypes as C
import os
import time

import numpy as np
import torch

from . import kernels as K
from .fno import FNO, Config
from .prior import green_1d

CIN = 6
# weights trained on 600 solver-generated sphere media (128^3 ROI, seeds 1000-1599)
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "spheres_128.npz")


class WarmStart:
    def __init__(self, cfg: Config | None = None, factor: int = 2, n_bg: float = 1.33, params=None,
                 seed: int = 0):
        self.cfg = cfg or Config(cin=CIN)
        assert self.cfg.cin == CIN
        self.factor = factor
        self.n_bg = n_bg
        self.net = FNO(self.cfg, params, seed=seed)
        self.dev = self.net.dev

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

    def upload_medium(self, sim, nmap_roi) -> torch.Tensor:
        """The ROI refractive index on the GPU in the network's [z][y][x] order
        (nmap_roi in [y][x][z] order, as the reference stores media).  Input
        staging, like the solver's upload of gamma."""
        n = torch.from_numpy(np.ascontiguousarray(np.real(nmap_roi).astype(np.float32).transpose(2, 0, 1)))
        return n.to(self.dev).reshape(-1)

    def _background(self, sim):
        """u(z) on the grid and on the ROI's planes, on the GPU (cached per geometry)."""
        key = (tuple(int(x) for x in sim.N), float(sim.k0), float(sim.grid.dx),
               tuple(int(x) for x in np.asarray(sim.roi).ravel()))
        if getattr(self, "_bg_key", None) != key:
            sl = self.roi_slices(sim)
            u_full = self.prior_u(sim)
            self._bg = (torch.from_numpy(u_full.view(np.float32).copy()).to(self.dev),
                        torch.from_numpy(u_full[sl[2]].view(np.float32).copy()).to(self.dev))
            self._bg_key = key
        return self._bg

    def prepare(self, sim, nmap_roi=None, n_dev=None) -> dict:
        """Device-side prior and network input from the ROI refractive index,
        given on the host (nmap_roi, [y][x][z]) or already on the GPU (n_dev,
        from upload_medium)."""
        sl = self.roi_slices(sim)
        nz, ny, nx = (sl[2].stop - sl[2].start, sl[0].stop - sl[0].start, sl[1].stop - sl[1].start)
        n = n_dev if n_dev is not None else self.upload_medium(sim, nmap_roi)
        u, u_roi = self._background(sim)
        phase = torch.empty(nz * ny * nx, dtype=torch.float32, device=self.dev)
        P = torch.empty(2 * nz * ny * nx, dtype=torch.float32, device=self.dev)
        shape = (C.c_int * 3)(nz, ny, nx)
        K.check(K.lib().wfno_wkb_prior(K.ptr(n), K.ptr(phase), K.ptr(P), K.ptr(u_roi), shape,
                                       float(sim.k0 * sim.grid.dx), float(self.n_bg), K.stream_ptr()))
        f = self.factor
        V = (nz // f) * (ny // f) * (nx // f)
        assert nz // f == self.cfg.L and ny // f == self.cfg.L and nx // f == self.cfg.L, "grid mismatch"
        X = torch.empty(CIN * V, dtype=torch.float32, device=self.dev)
        K.check(K.lib().wfno_features(K.ptr(n), K.ptr(phase), K.ptr(X), shape, f, float(self.n_bg),
                                      K.stream_ptr()))
        return dict(X=X, P=P, u=u, shape=(nz, ny, nx))

    def _pool(self, a):
        f = self.factor
        Z, Y, X = a.shape
        return a.reshape(Z // f, f, Y // f, f, X // f, f).mean(axis=(1, 3, 5))

    def target(self, sim, E_roi, prep) -> np.ndarray:
        """Envelope of the solved field E_roi ([y][x][z] complex, ROI) relative to the prior."""
        nz, ny, nx = prep["shape"]
        P = prep["P"].cpu().numpy().view(np.complex64).reshape(nz, ny, nx)
        A = E_roi.transpose(2, 0, 1) / P - 1
        A = self._pool(A)
        return np.stack([A.real, A.imag]).reshape(2, -1).astype(np.float32)

    # -- inference -----------------------------------------------------------------
    def forward(self, X: torch.Tensor, graph: bool = True) -> torch.Tensor:
        """Network output [1][2][L^3] for the input X ([CIN][L^3]).  With
        graph=True the ~50 kernel launches of one forward run as one replay of
        a captured HIP graph (built on first use; the weights are read in
        place, so training between calls is seen by later replays)."""
        if not graph:
            y, _ = self.net.forward(X.reshape(1, CIN, -1), keep=False)
            return y
        if getattr(self, "_graph", None) is None:
            self._gx = X.detach().clone()
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                for _ in range(2):  # allocates the cached work buffers outside the capture
                    self.net.forward(self._gx.reshape(1, CIN, -1), keep=False)
            torch.cuda.current_stream().wait_stream(side)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._gy, _ = self.net.forward(self._gx.reshape(1, CIN, -1), keep=False)
            self._graph = g
        self._gx.copy_(X.reshape(self._gx.shape))
        self._graph.replay()
        return self._gy

    def predict_E0(self, sim, prep, graph: bool = True) -> torch.Tensor:
        """Warm-start field on the solver grid (device, solver layout)."""
        return self.compose(sim, prep, self.forward(prep["X"], graph))

    def compose(self, sim, prep, y: torch.Tensor) -> torch.Tensor:
        E0 = torch.empty(2 * int(np.prod(sim.N[:3])), dtype=torch.float32, device=self.dev)
        sl = self.roi_slices(sim)
        grid = (C.c_int * 3)(*sim.N[:3])
        lo = (C.c_int * 3)(*[s.start for s in sl])
        n = (C.c_int * 3)(*[s.stop - s.start for s in sl])
        foot = (C.c_int * 4)(sl[0].start, sl[0].stop, sl[1].start, sl[1].stop)
        K.check(K.lib().wfno_compose_prior(K.ptr(E0), K.ptr(y), K.ptr(prep["P"]), K.ptr(prep["u"]), grid, lo, n,
                                           self.factor, foot, K.stream_ptr()))
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
        prep = ws.prepare(sim, nmap)
        X.append(prep["X"].cpu().numpy().reshape(CIN, -1))
        T.append(ws.target(sim, E, prep))
        info.append((seed, r.iterations, dt))
        if log and len(info) % 50 == 0:
            log(f"generated {len(info)} samples (last: {r.iterations} iterations, {dt * 1e3:.1f} ms)")
    return np.stack(X), np.stack(T), info, holder


def evaluate(ws: WarmStart, seeds, holder=None, n=128, boundary=32, tight=1e-10):
    """Per seed: zero start, prior-only warm start and learned warm start, each
    run to the zero start's absolute residual (same stop denominator), timed
    from inputs resident on the GPU; errors against a tight zero-start solve."""
    rows = []
    # build the forward graph before any clock runs
    ws.forward(torch.zeros(CIN * ws.cfg.L ** 3, dtype=torch.float32, device=ws.dev))
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
        n_dev = ws.upload_medium(sim, nmap)  # input staging, outside the clock like gamma's upload
        ws._background(sim)
        for name in ("prior", "learned"):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            prep = ws.prepare(sim, n_dev=n_dev)
            if name == "learned":
                E0 = ws.predict_E0(sim, prep)
            else:
                zero = torch.zeros(2 * ws.cfg.L ** 3, dtype=torch.float32, device=ws.net.dev)
                E0 = ws.compose(sim, prep, zero)
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
