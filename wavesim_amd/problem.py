"""Problem set-up through the library's C++ implementation
(``wsim_problem_query`` / ``wsim_problem_build``): the WaveSim / WaveSimVector
constructor (WaveSimBase.m:40-82) for callers that do not run the Python
mirror in ``simulation.py``; both give the same gamma, epsilon, k02e and
wiggle tables."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import capi


def build(refractive_index, options: dict, vector: bool = False):
    """Returns dict(N, roi, Bl, Br, gamma, wiggles, epsilon, k02e, k0c,
    iterations_per_cycle) for the options the Python mirror takes
    (pixel_size, lambda, boundary_widths, epsilon, epsilonmin, ACC, roi)."""
    L = capi.lib()
    n = np.asarray(refractive_index)
    if n.ndim == 1:
        n = n[:, None]
    if n.ndim > 3:
        raise ValueError("refractive index must be 1-D, 2-D or 3-D")
    shape = list(n.shape) + [1] * (3 - n.ndim)
    d = capi.problem_desc_t()
    for a in range(3):
        d.shape[a] = shape[a]
    d.pixel_size = float(options["pixel_size"])
    d.lambda_ = float(options.get("lambda", 1.0))
    bw = [int(b) for b in np.ravel(options["boundary_widths"])] + [0, 0, 0]
    for a in range(3):
        d.boundary_widths[a] = bw[a]
    d.epsilon = float(options["epsilon"]) if options.get("epsilon") is not None else 0.0
    d.epsilonmin = float(options.get("epsilonmin", 3))
    acc = np.ravel(options.get("wiggle", options.get("ACC", True)))
    if acc.size == 1:
        flags = [-1, -1, -1] if bool(acc[0]) else [0, 0, 0]
    else:
        flags = [1 if bool(x) else 0 for x in acc]
    for a in range(3):
        d.acc[a] = flags[a]
    d.vector = 1 if vector else 0
    if options.get("roi") is not None:
        r = np.asarray(options["roi"])
        d.has_roi = 1
        for i in range(2):
            v = [int(x) for x in np.ravel(r[i])] + [1] * 4
            for a in range(4):
                d.roi[i][a] = v[a]
    info = capi.problem_info_t()
    capi.check(L.wsim_problem_query(C.byref(d), C.byref(info)))
    N = [int(info.N[a]) for a in range(4)]
    nmap = np.asfortranarray(n.reshape(shape).astype(np.complex128))
    gamma = np.empty(N[:3], dtype=np.complex64, order="F")
    tables, keep = [], []
    wig = (capi.wiggle_t * info.nwiggles)()
    for w in range(info.nwiggles):
        t = dict(pxe=np.empty(N[1], np.float32), pye=np.empty(N[0], np.float32), pze=np.empty(N[2], np.float32),
                 gx=np.empty(N[1], np.complex64), gy=np.empty(N[0], np.complex64), gz=np.empty(N[2], np.complex64))
        tables.append(t)
        for k, v in t.items():
            setattr(wig[w], k, v.ctypes.data)
    capi.check(L.wsim_problem_build(C.byref(d), nmap.ctypes.data, C.byref(info), gamma.ctypes.data, wig))
    return dict(N=N, roi=np.array([[int(info.roi[i][a]) for a in range(4)] for i in range(2)]),
                Bl=[int(x) for x in info.Bl], Br=[int(x) for x in info.Br], gamma=gamma, wiggles=tables,
                epsilon=np.float32(info.epsilon), k02e=np.complex64(complex(info.k02e.re, info.k02e.im)),
                k0c=float(info.k0c), iterations_per_cycle=float(info.iterations_per_cycle))
