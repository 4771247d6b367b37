"""FFT plan cache for the autotuner.

A plan (per-axis Stockham radix order and workgroup tile width, see
``wsim_solver_set_plan``) is keyed by grid size, components, right-hand sides
and wiggle count.  Lookups consult the user cache (``$WSIM_PLAN_CACHE``,
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


def key(N, nrhs: int, nw: int) -> str:
    return "x".join(str(int(n)) for n in N) + f"|rhs{int(nrhs)}|nw{int(nw)}"


def _load(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(N, nrhs: int, nw: int) -> Optional[str]:
    k = key(N, nrhs, nw)
    for path in (cache_path(), SHIPPED):
        d = _load(path)
        if k in d:
            return d[k]["plan"] if isinstance(d[k], dict) else d[k]
    return None


def store(N, nrhs: int, nw: int, plan: str, ms: float) -> None:
    path = cache_path()
    d = _load(path)
    d[key(N, nrhs, nw)] = {"plan": plan, "ms_per_iteration": ms}
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
    except OSError:
        pass
