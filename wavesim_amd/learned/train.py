"""Generate solver data, train the warm-start network on the GPU, evaluate A/B.

    python -m wavesim_amd.learned.train --samples 400 --epochs 12 --out warmstart.npz

Prints progress lines and a final JSON summary (zero-start vs learned warm
start time-to-residual on held-out media).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from .fno import Config
from .warmstart import WarmStart, evaluate, generate


def log(msg):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=400)
    ap.add_argument("--eval", type=int, default=8)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--boundary", type=int, default=32)
    ap.add_argument("--width", type=int, default=32)
    ap.add_argument("--modes", type=int, default=16)
    ap.add_argument("--layers", type=int, default=4)
    ap.add_argument("--factor", type=int, default=2)
    ap.add_argument("--out", default="")
    ap.add_argument("--load", default="")
    a = ap.parse_args(argv)
    L = a.n // a.factor
    if a.load:
        ws = WarmStart.load(a.load)
    else:
        ws = WarmStart(Config(L=L, cin=6, width=a.width, modes=a.modes, layers=a.layers, proj=64),
                       factor=a.factor)
    log(f"network: {ws.net.nparam} parameters, grid {L}^3")
    summary = {"parameters": ws.net.nparam, "grid": L}
    holder = None
    if a.samples > 0:
        t0 = time.perf_counter()
        X, T, info, holder = generate(ws, range(1000, 1000 + a.samples), a.n, a.boundary, log=log)
        its = np.array([i[1] for i in info])
        summary["data"] = dict(samples=len(info), seconds=round(time.perf_counter() - t0, 1),
                               mean_iterations=float(its.mean()),
                               mean_solve_ms=float(np.mean([i[2] for i in info]) * 1e3))
        log(f"data: {summary['data']}")
        Xd = torch.from_numpy(X).cuda()
        Td = torch.from_numpy(T).cuda()
        t0 = time.perf_counter()
        hist = ws.train(Xd, Td, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log)
        torch.cuda.synchronize()
        summary["train"] = dict(epochs=a.epochs, seconds=round(time.perf_counter() - t0, 1),
                                loss_first=hist[0], loss_last=hist[-1],
                                target_mean_square=float((T ** 2).mean()))
        log(f"train: {summary['train']}")
        del Xd, Td
        if a.out:
            ws.save(a.out)
    rows, holder = evaluate(ws, range(a.eval), holder, a.n, a.boundary)
    for r in rows:
        log(json.dumps(r))
    zs = np.array([r["zero_s"] for r in rows])
    ls = np.array([r["learned_s"] for r in rows])
    ps = np.array([r["prior_s"] for r in rows])
    summary["eval"] = dict(
        media=len(rows),
        zero_its=float(np.mean([r["zero_its"] for r in rows])),
        prior_its=float(np.mean([r["prior_its"] for r in rows])),
        learned_its=float(np.mean([r["learned_its"] for r in rows])),
        zero_ms=float(zs.mean() * 1e3), prior_ms=float(ps.mean() * 1e3), learned_ms=float(ls.mean() * 1e3),
        speedup_learned=float(zs.sum() / ls.sum()), speedup_prior=float(zs.sum() / ps.sum()),
        zero_err=float(np.mean([r["zero_err"] for r in rows])),
        learned_err=float(np.mean([r["learned_err"] for r in rows])),
        prior_err=float(np.mean([r["prior_err"] for r in rows])))
    print(json.dumps(summary), flush=True)
    return summary


if __name__ == "__main__":
    main()
