"""Accuracy of the GPU solver per benchmark configuration (run on a GPU host).

    python -m tests.accuracy_report cfg1 cfg2 cfg3b cfg5     # or cfg3a alone (slow oracle)

For each configuration the GPU (complex64) result is compared with
* the oracle in complex64 (the reference algorithm restated, same inputs):
  the parity figure; iteration counts are reported side by side;
* the oracle in complex128 (single_precision = false): the spread that
  single precision itself introduces;
* the analytic solution where one exists (cfg1: (i h^2/4) H0(k r) of a point
  source in 2-D, on an annulus away from the source and the boundary).
Prints one JSON line per configuration; oracle progress lines go to stderr.
This is a report, not a test: the oracle runs at full size on the host CPU.
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def _rel(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _configs():
    from wavesim_amd import workloads as W
    return {
        "cfg1": ("2-D homogeneous 256^2 + 32 px, point source", lambda: W.homogeneous_2d(256, 32), False),
        "cfg2": ("2-D disordered 1024^2 + 64 px, plane wave", lambda: W.disordered_2d(1024, 64), False),
        "cfg3a": ("3-D disordered 256^3 + 32 px, point source", lambda: W.disordered_3d(256, 32), False),
        "cfg3b": ("3-D vector, sphere n 1.6 in 1.33, 128^3 + 32 px, x-polarised plane wave",
                  lambda: W.sphere_vector_3d(128, 32), True),
        "cfg5": ("3-D spheres 128^3 + 32 px (learned warm-start workload), plane wave",
                 lambda: W.spheres_3d(128, 32, seed=0), False),
    }


def _oracle_source(src):
    from oracle import mbs
    s = mbs.Source()
    for p, v in zip(src.positions, src.values):
        s = s + mbs.Source.point(np.asarray(v), list(p))
    return s


def run(name):
    import wavesim_amd as w
    from oracle import analytic, mbs
    label, build, vector = _configs()[name]
    nmap, opt, src = build()
    sim = (w.WaveSimVector if vector else w.WaveSim)(nmap, opt)
    t0 = time.perf_counter()
    Eg, sg = sim.exec(src)
    t_gpu = time.perf_counter() - t0
    out = {"config": name, "label": label, "grid": [int(x) for x in sim.N], "gpu_iterations": sg.iterations,
           "gpu_s": round(t_gpu, 4)}
    workers = os.cpu_count()
    for prec, single in (("c64", True), ("c128", False)):
        o = mbs.make_sim(nmap, dict(opt, single_precision=single), vector=vector)
        o.callback_interval = 40  # progress only: the energy cadence was fixed at construction

        def progress(_sim, st, prec=prec):
            print(f"# {name} oracle {prec}: iteration {st.it - 1}, {time.perf_counter() - t1:.0f} s",
                  file=sys.stderr, flush=True)

        t1 = time.perf_counter()
        Eo, so = mbs.run(o, _oracle_source(src), workers=workers, callback=progress)
        out[f"oracle_{prec}_iterations"] = so.it - 1
        out[f"oracle_{prec}_s"] = round(time.perf_counter() - t1, 1)
        out[f"rel_l2_gpu_vs_oracle_{prec}"] = _rel(Eg, Eo)
        if prec == "c128":
            Eo128 = Eo
        else:
            Eo64 = Eo
    out["rel_l2_oracle_c64_vs_c128"] = _rel(Eo64, Eo128)
    if name == "cfg1":
        n = nmap.shape[0]
        c = n // 2
        yy, xx = np.meshgrid(np.arange(n) - (c - 1), np.arange(n) - (c - 1), indexing="ij")
        r = np.hypot(yy, xx) * opt["pixel_size"]
        ref = analytic.hankel_2d(sim.k0, opt["pixel_size"], np.where(r > 0, r, 1))
        mask = (r > 2.0) & (r < 0.45 * n * opt["pixel_size"])
        out["rel_l2_gpu_vs_analytic_annulus"] = _rel(Eg[..., 0, 0][mask] if Eg.ndim == 4 else Eg[..., 0][mask],
                                                     ref[mask])
        out["rel_l2_oracle_c128_vs_analytic_annulus"] = _rel(Eo128[..., 0, 0][mask], ref[mask])
    print(json.dumps(out), flush=True)
    return out


def main(argv=None):
    names = (argv if argv is not None else sys.argv[1:]) or ["cfg1", "cfg2", "cfg3b", "cfg5"]
    for n in names:
        run(n)


if __name__ == "__main__":
    main()
