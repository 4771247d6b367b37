"""ctypes binding of ``libwavesim_fno.so`` (include/wavesim_fno.h).

Device memory is passed as raw pointers (torch tensors are used only to own
device memory and streams); every call launches one of the project's HIP
kernels.  The library is required -- there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("WFNO_LIB", os.path.join(_HERE, "libwavesim_fno.so"))


class gemm_args(C.Structure):
    _fields_ = [("M", C.c_longlong), ("N", C.c_longlong), ("K", C.c_longlong),
                ("nb1", C.c_int), ("nb2", C.c_int), ("nb3", C.c_int),
                ("A", C.c_void_p), ("sam", C.c_longlong), ("sak", C.c_longlong),
                ("sab1", C.c_longlong), ("sab2", C.c_longlong), ("sab3", C.c_longlong),
                ("B", C.c_void_p), ("sbk", C.c_longlong), ("sbn", C.c_longlong),
                ("sbb1", C.c_longlong), ("sbb2", C.c_longlong), ("sbb3", C.c_longlong),
                ("C", C.c_void_p), ("scm", C.c_longlong), ("scn", C.c_longlong),
                ("scb1", C.c_longlong), ("scb2", C.c_longlong), ("scb3", C.c_longlong),
                ("alpha", C.c_float), ("beta", C.c_float)]


P = C.c_void_p
LL = C.c_longlong
SYMBOLS = {
    "wfno_gemm": [C.POINTER(gemm_args), P],
    "wfno_bias_act": [P, P, P, LL, LL, LL, C.c_int, P],
    "wfno_gelu_bwd": [P, P, P, LL, P],
    "wfno_bias_grad": [P, P, LL, LL, LL, C.c_int, P],
    "wfno_mse": [P, P, P, P, C.c_int, LL, C.c_float, P],
    "wfno_reduce_chunks": [P, P, C.c_int, LL, C.c_int, P],
    "wfno_complex_weights": [P, P, LL, C.c_int, LL, C.c_int, C.c_int, C.c_int, P],
    "wfno_wkb_prior": [P, P, P, P, C.POINTER(C.c_int), C.c_float, C.c_float, P],
    "wfno_features": [P, P, P, C.POINTER(C.c_int), C.c_int, C.c_float, P],
    "wfno_compose_prior": [P, P, P, P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int,
                           C.POINTER(C.c_int), P],
    "wfno_adam": [P, P, P, P, LL, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, P],
    "wfno_compose_warm": [P, P, P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int,
                          C.POINTER(C.c_int), P],
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m wavesim_amd.build`")
        L = C.CDLL(LIB_PATH)
        for n, args in SYMBOLS.items():
            f = getattr(L, n)
            f.restype = C.c_int
            f.argtypes = args
        L.wfno_last_error.restype = C.c_char_p
        L.wfno_last_error.argtypes = []
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise RuntimeError(f"wavesim_fno error {rc}: {lib().wfno_last_error().decode()}")


def ptr(t) -> int:
    return int(t.data_ptr())


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _addr(x) -> int:
    if isinstance(x, tuple):
        t, off = x
        return ptr(t) + 4 * int(off)
    return ptr(x)


def gemm(M, N, K, A, sA, B, sB, Cm, sC, alpha=1.0, beta=0.0, nb=(1, 1, 1), stream=None):
    """C = alpha A B + beta C for a batch nb = (nb1, nb2, nb3) with element
    strides sA = (sam, sak, sab1, sab2, sab3), sB = (sbk, sbn, sbb1, sbb2, sbb3),
    sC = (scm, scn, scb1, scb2, scb3).  A, B, C are float32 device tensors or
    (tensor, element offset) pairs."""
    nb = tuple(nb) + (1,) * (3 - len(nb))
    sA = tuple(sA) + (0,) * (5 - len(sA))
    sB = tuple(sB) + (0,) * (5 - len(sB))
    sC = tuple(sC) + (0,) * (5 - len(sC))
    g = gemm_args(int(M), int(N), int(K), int(nb[0]), int(nb[1]), int(nb[2]),
                  _addr(A), *[int(s) for s in sA], _addr(B), *[int(s) for s in sB],
                  _addr(Cm), *[int(s) for s in sC], float(alpha), float(beta))
    check(lib().wfno_gemm(C.byref(g), stream_ptr(stream)))
