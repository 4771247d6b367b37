"""Writes the full-size golden fixtures of this directory from the oracle
(TEST INFRASTRUCTURE ONLY): the complex64 oracle solve of the 256^3
disordered workload (320^3 grid) and of the 128^3 vector sphere workload
(192^3 x 3 grid), run to convergence on this host's cores.  Whole fields are
too large to commit, so each fixture keeps what a GPU test can compare at
full size without running the oracle: the iteration count and diff_energy
history, the ROI field sampled every `STEP`-th voxel along each axis, and two
projections of the whole ROI field on seeded random unit phases (every voxel
contributes).  Takes about 10 minutes on 8 cores.

    python -m tests.golden.make_fullsize_fixtures
"""
from __future__ import annotations

import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
STEP = {"fullsize_disordered_3d_256": 8, "fullsize_sphere_vector_3d_128": 6}


def projections(E, seed=7, k=2):
    """k sums of E times seeded random unit phases over the whole array (complex128)."""
    rng = np.random.default_rng(seed)
    E = np.asarray(E, np.complex128)
    out = []
    for _ in range(k):
        w = np.exp(2j * np.pi * rng.random(E.shape))
        out.append(np.sum(w * E))
    return np.array(out)


def main():
    from oracle import mbs
    from wavesim_amd.workloads import WORKLOADS
    for name, wl, vector in (("fullsize_disordered_3d_256", "disordered_3d_256", False),
                             ("fullsize_sphere_vector_3d_128", "sphere_vector_3d_128", True)):
        nmap, opt, src = WORKLOADS[wl]()
        o = mbs.make_sim(nmap, opt, vector=vector)
        s = mbs.Source(list(src.positions), [np.asarray(v) for v in src.values])
        t0 = time.perf_counter()
        E, st = mbs.run(o, s, workers=os.cpu_count())
        k = STEP[name]
        np.savez_compressed(os.path.join(HERE, name + ".npz"), workload=np.array(wl), vector=np.array(vector),
                            iterations=np.array(st.it - 1), converged=np.array(bool(st.converged)),
                            diff_energy=np.asarray(st.diff_energy, np.float64), step=np.array(k),
                            sample=E[::k, ::k, ::k].astype(np.complex64),
                            proj=projections(E), norm=np.array(np.linalg.norm(E.astype(np.complex128))))
        print(f"{name}: {st.it - 1} iterations, {time.perf_counter() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main()
