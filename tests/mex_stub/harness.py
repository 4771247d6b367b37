"""ctypes side of the fake MATLAB API (tests/mex_stub/mxfake.cpp): builds the
12 inputs WaveSimBase.run_algorithm hands to RunWaveSimAlgorithmMex
(WaveSimBase.m:145-146) from a simulation of the oracle, calls mexFunction
and reads the outputs and the recorded SimulationCallback calls.  Test
infrastructure only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libmexshim.so")


class Mex:
    def __init__(self, path: str = LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is missing: build it with tests/mex_stub/build_shim.py")
        L = C.CDLL(path)
        P = C.c_void_p
        sig = {
            "mxf_numeric": (P, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64), P]),
            "mxf_string": (P, [C.c_char_p]),
            "mxf_cell": (P, [C.c_int64]),
            "mxf_cell_set": (None, [P, C.c_int64, P]),
            "mxf_struct": (P, []),
            "mxf_object": (P, []),
            "mxf_set": (None, [P, C.c_char_p, P]),
            "mxf_destroy": (None, [P]),
            "mxf_numel": (C.c_int64, [P]),
            "mxf_ndims": (C.c_int, [P]),
            "mxf_dim": (C.c_int64, [P, C.c_int]),
            "mxf_class": (C.c_int, [P]),
            "mxf_is_complex": (C.c_int, [P]),
            "mxf_data": (P, [P]),
            "mxf_live_temporaries": (C.c_int64, []),
            "mxf_call": (C.c_int, [C.c_int, C.POINTER(P), C.c_int, C.POINTER(P), C.c_char_p, C.c_int]),
            "mxf_log_clear": (None, []),
            "mxf_log_count": (C.c_int, []),
            "mxf_log_entry": (C.c_int, [C.c_int, C.c_char_p, C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_double),
                                        C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                        C.POINTER(C.c_double), C.POINTER(C.c_int)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        self.L = L

    # -- arrays (MATLAB index order in, column-major storage) -------------------------
    def array(self, a, single: bool):
        a = np.asarray(a)
        cplx = np.iscomplexobj(a)
        dt = (np.complex64 if cplx else np.float32) if single else (np.complex128 if cplx else np.float64)
        f = np.asfortranarray(a.astype(dt))
        shape = list(a.shape)
        while len(shape) > 2 and shape[-1] == 1:  # MATLAB drops trailing singleton dimensions
            shape.pop()
        while len(shape) < 2:
            shape.append(1)
        dims = (C.c_int64 * len(shape))(*shape)
        return self.L.mxf_numeric(1 if single else 0, 1 if cplx else 0, len(shape), dims, f.ctypes.data)

    def cell(self, items):
        c = self.L.mxf_cell(len(items))
        for i, v in enumerate(items):
            self.L.mxf_cell_set(c, i, v)
        return c

    def struct(self, fields, obj=False):
        s = self.L.mxf_object() if obj else self.L.mxf_struct()
        for k, v in fields.items():
            self.L.mxf_set(s, k.encode(), v)
        return s

    def read(self, a) -> np.ndarray:
        dims = [self.L.mxf_dim(a, k) for k in range(self.L.mxf_ndims(a))]
        single = self.L.mxf_class(a) == 1
        cplx = self.L.mxf_is_complex(a) == 1
        dt = (np.complex64 if cplx else np.float32) if single else (np.complex128 if cplx else np.float64)
        n = int(np.prod(dims))
        buf = (C.c_char * (n * np.dtype(dt).itemsize)).from_address(self.L.mxf_data(a))
        return np.frombuffer(buf, dtype=dt).copy().reshape(dims, order="F")

    # -- the call of WaveSimBase.m:145-146 ----------------------------------------------
    def inputs(self, sim, source, max_iter=np.inf, callback="NoCallback", interval=50.0):
        """sim: oracle.mbs.Sim; source: the prepared oracle Source (grid positions)."""
        N = np.asarray(sim.N, np.float64).reshape(1, 4)
        wig = []
        for w in sim.wiggles:
            Nx, Ny, Nz = len(w["pxe"]), len(w["pye"]), len(w["pze"])
            wig.append(self.struct({
                "pxe": self.array(w["pxe"].reshape(1, Nx), True), "pye": self.array(w["pye"].reshape(Ny, 1), True),
                "pze": self.array(w["pze"].reshape(1, 1, Nz), True),
                "gx": self.array(w["gx"].reshape(1, Nx), True), "gy": self.array(w["gy"].reshape(Ny, 1), True),
                "gz": self.array(w["gz"].reshape(1, 1, Nz), True)}))
        src = self.struct({
            "positions": self.cell([self.array(np.asarray(p, np.float64).reshape(1, 4), False)
                                    for p in source.positions]),
            "values": self.cell([self.array(v, True) for v in source.values])}, obj=True)
        return [self.array(N, False), self.array(sim.gamma, True),
                self.array(np.float64(len(sim.wiggles)).reshape(1, 1), False),
                self.cell(wig), src, self.array(np.float32(sim.epsilon).reshape(1, 1), True),
                self.array(np.complex64(sim.k02e).reshape(1, 1), True),
                self.array(np.asarray(sim.roi, np.float64), False),
                self.array(np.float64(max_iter).reshape(1, 1), False),
                self.array(np.float64(sim.energy_threshold).reshape(1, 1), False),
                self.L.mxf_string(callback.encode()), self.array(np.float64(interval).reshape(1, 1), False)]

    def call(self, prhs, nlhs=2):
        P = C.c_void_p
        out = (P * nlhs)()
        arr = (P * len(prhs))(*prhs)
        err = C.create_string_buffer(512)
        rc = self.L.mxf_call(nlhs, out, len(prhs), arr, err, 512)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        res = [self.read(out[i]) for i in range(nlhs)]
        for i in range(nlhs):
            self.L.mxf_destroy(out[i])
        return res

    def free(self, prhs):
        for a in prhs:
            self.L.mxf_destroy(a)

    def log(self):
        out = []
        for i in range(self.L.mxf_log_count()):
            name, typ = C.create_string_buffer(64), C.create_string_buffer(64)
            de = (C.c_double * 100000)()
            nde, nen, thr, nargs = C.c_int64(), C.c_int64(), C.c_double(), C.c_int()
            self.L.mxf_log_entry(i, name, 64, typ, 64, de, 100000, C.byref(nde), C.byref(nen), C.byref(thr),
                                 C.byref(nargs))
            out.append(dict(name=name.value.decode(), type=typ.value.decode(), de=np.asarray(de[:nde.value]),
                            energy_numel=nen.value, threshold=thr.value, nargs=nargs.value))
        return out
