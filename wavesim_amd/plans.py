"""FFT plan cache for the autotuner.

A plan (per-axis Stockham radix order and workgroup tile width, see
``wsim_solver_set_plan``) is keyed by grid size, components, right-hand sides,
wiggle count and precision (complex128 solves, whose elements take twice the
registers and LDS, have plans of their own).  Lookups consult the user cache (``$WSIM_PLAN_CACHE``,
default ``~/.cache/wavesim_amd/plans.json``) and then the plans shipped with
the package (``tuned_plans.json``, produced by ``MBSSolver.tune`` on MI355X).
"""
from __future__ import annotations

import json
import os
from typing import Optional

_HERE = os.path.dirname(os.path.abspath(__file__))
SHIPPED = os.path.join(_HERE, "tuned_plans.json")


def cache_path() -> str:
    return os.environ.get("WSIM_PLAN_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "wavesim_amd", "plans.json"))


def key(N, nrhs: int, nw: int, f64: bool = False) -> str:
    return "x".join(str(int(n)) for n in N) + f"|rhs{int(nrhs)}|nw{int(nw)}" + ("|f64" if f64 else "")


def _load(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(N, nrhs: int, nw: int, f64: bool = False) -> Optional[str]:
    """The plan for the shape; a complex128 solve without a plan of its own
    takes the complex64 one (the library's shipped table does the same)."""
    keys = [key(N, nrhs, nw, f64)] + ([key(N, nrhs, nw)] if f64 else [])
    tables = [_load(cache_path()), _load(SHIPPED)]
    for k in keys:
        for d in tables:
            if k in d:
                return d[k]["plan"] if isinstance(d[k], dict) else d[k]
    return None


def store(N, nrhs: int, nw: int, plan: str, ms: float, f64: bool = False) -> None:
    path = cache_path()
    d = _load(path)
    d[key(N, nrhs, nw, f64)] = {"plan": plan, "ms_per_iteration": ms}
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
    except OSError:
        pass
