"""Writes the golden fixtures of this directory from the oracle (TEST
INFRASTRUCTURE ONLY): the oracle's complex64 solves and propagations of small
seeded problems, with their inputs.  The reference holds no golden vectors of
its own (its .mat outputs are git-ignored), so these pin the oracle's current
behaviour and give the GPU tests a fixed anchor that needs no oracle run.

    python -m tests.golden.make_fixtures
"""
from __future__ import annotations

import json
import os

import numpy as np

from tests import cases

HERE = os.path.dirname(os.path.abspath(__file__))

SOLVES = {
    "solve_2d_disordered": (cases.case_2d_disordered, False),
    "solve_2d_homogeneous": (cases.case_2d_homogeneous, False),
    "solve_1d": (cases.case_1d, False),
    "solve_periodic_2d": (cases.case_periodic_2d, False),
    "solve_3d_disordered": (cases.case_3d_disordered, False),
    "solve_3d_vector": (cases.case_3d_vector_sphere, True),
}


def pack_sources(src):
    out = {"n_src": np.array(len(src))}
    for k, (v, p) in enumerate(src):
        out[f"src_val_{k}"] = np.asarray(v, dtype=np.complex128)
        out[f"src_pos_{k}"] = np.asarray(p, dtype=np.int64)
    return out


def unpack_sources(d):
    return [(d[f"src_val_{k}"] if d[f"src_val_{k}"].ndim else complex(d[f"src_val_{k}"]),
             [int(x) for x in d[f"src_pos_{k}"]]) for k in range(int(d["n_src"]))]


def load(name):
    """(nmap, opt, sources, vector, arrays) of one fixture."""
    d = np.load(os.path.join(HERE, name + ".npz"))
    opt = json.loads(str(d["opt_json"]))
    return d["nmap"], opt, unpack_sources(d), bool(d["vector"]), d


def main():
    from oracle import mbs
    for name, (builder, vector) in SOLVES.items():
        nmap, opt, src = builder()
        o = mbs.make_sim(nmap, opt, vector=vector)
        E, st = mbs.run(o, cases.oracle_source(src))
        arrays = dict(nmap=np.asarray(nmap), opt_json=np.array(json.dumps(opt)), vector=np.array(vector),
                      E=E.astype(np.complex64), iterations=np.array(st.it - 1),
                      converged=np.array(bool(st.converged)),
                      diff_energy=np.asarray(st.diff_energy, dtype=np.float64), **pack_sources(src))
        if name in ("solve_2d_disordered", "solve_3d_disordered", "solve_3d_vector"):
            rng = np.random.default_rng(11)
            Ein = (rng.standard_normal(o.N) + 1j * rng.standard_normal(o.N)).astype(np.complex64)
            arrays["prop_in"] = Ein
            ws = sorted({0, len(o.wiggles) - 1})  # the first and the last wiggle (size)
            for w in ws:
                arrays[f"prop_out_{w}"] = np.asarray(mbs.propagate(o, Ein, o.wiggles[w]), dtype=np.complex64)
            arrays["prop_wiggles"] = np.array(ws)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print(name, E.shape, st.it - 1, flush=True)


if __name__ == "__main__":
    main()
