"""ctypes driver of the CPU oracle (test infrastructure only).

Loads oracle/liboracle.so (built by oracle/Makefile) and exposes the same
methods as xmd_amd.engine.Engine, so a parity test runs identical calls on
both.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
import this module.
"""
import ctypes as C
import os

import numpy as np

from xmd_amd import _abi
from xmd_amd._host import HostSystem

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_LIBS = {}


def load(path):
    """The oracle library at `path` (loaded once per path)."""
    if path not in _LIBS:
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle` or __graft_entry__.build()")
        L = C.CDLL(path)
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_rng_next.restype = C.c_uint64
        L.oracle_rng_next.argtypes = [C.POINTER(C.c_uint64)]
        L.oracle_rng_uniformf.restype = C.c_float
        L.oracle_rng_uniformf.argtypes = [C.POINTER(C.c_uint64)]
        L.oracle_rand_normals.argtypes = [C.c_uint64, C.c_int64, C.POINTER(C.c_float), C.POINTER(C.c_uint64)]
        L.oracle_prim.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.oracle_create.argtypes = [C.POINTER(_abi.XmdParams), C.POINTER(_abi.XmdTopology), C.c_int, C.c_void_p]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_step.argtypes = [C.c_void_p, C.c_int64, C.c_int]
        _LIBS[path] = L
    return _LIBS[path]


def lib():
    return load(LIB_PATH)


class Oracle(HostSystem):
    """FP64 (real_kind=0) or float (real_kind=1) CPU oracle of a system;
    lib_path selects another build of the oracle (the mutant builds of the
    value-pin tests)."""

    def __init__(self, params, system, real_kind=0, lib_path=None):
        super().__init__(params, system)
        self._lib = load(lib_path or LIB_PATH)
        h = C.c_void_p()
        rc = self._lib.oracle_create(C.byref(params), C.byref(self.topo), int(real_kind), C.byref(h))
        if rc != 0:
            raise RuntimeError(self._lib.oracle_last_error().decode())
        self.handle = h

    def _fn(self, name):
        return getattr(self._lib, "oracle_" + name)

    def _err(self):
        return self._lib.oracle_last_error().decode()

    def step(self, nsteps, nthreads=0):
        self._check(self._lib.oracle_step(self.handle, C.c_int64(nsteps), int(nthreads)))

    def close(self):
        if self.handle:
            self._lib.oracle_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rng_next(state):
    s = C.c_uint64(state)
    v = lib().oracle_rng_next(C.byref(s))
    return v, s.value


def rand_normals(state, n):
    out = np.zeros(3 * n, dtype=np.float32)
    so = C.c_uint64()
    lib().oracle_rand_normals(C.c_uint64(state), C.c_int64(n), out.ctypes.data_as(C.POINTER(C.c_float)),
                              C.byref(so))
    return out.reshape(n, 3), so.value


def morph_into_saw(state, chain_lens, orig=None, bond=None, density=1.0, infer_box=False):
    """Positions (n, 3), box and generator state after morphing chains of the
    given lengths (model.cpp:92-174)."""
    n = int(sum(chain_lens))
    lens = np.asarray(chain_lens, dtype=np.int32)
    o = np.zeros((n, 3)) if orig is None else np.ascontiguousarray(orig, dtype=np.float64)
    pos = np.zeros((n, 3))
    box = np.zeros(3)
    s = C.c_uint64(state)
    lib().oracle_morph_into_saw(C.byref(s), C.c_int(len(lens)), lens.ctypes.data_as(C.POINTER(C.c_int)),
                                o.ctypes.data_as(C.POINTER(C.c_double)), C.c_double(-1.0 if bond is None else bond),
                                C.c_double(density), C.c_int(int(infer_box)),
                                pos.ctypes.data_as(C.POINTER(C.c_double)), box.ctypes.data_as(C.POINTER(C.c_double)))
    return pos, box, s.value


def prim(kind, *args):
    a = np.array(args, dtype=np.float64)
    out = np.zeros(2)
    lib().oracle_prim(int(kind), a.ctypes.data_as(C.POINTER(C.c_double)), out.ctypes.data_as(C.POINTER(C.c_double)))
    return out[0], out[1]
