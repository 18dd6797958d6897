"""ctypes driver of libxmdhip.so (the HIP engine, include/xmd_hip.h).

The library must be built (``__graft_entry__.build()`` or
``python -m xmd_amd.build``); there is no fallback path: a missing library or
a missing GPU raises.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from ._host import HostSystem, XmdError

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxmdhip.so")
_LIB = None


def lib():
    """Load the engine library (raises if it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.xmd_last_error.restype = C.c_char_p
        L.xmd_last_error.argtypes = [C.c_void_p]
        L.xmd_version.restype = C.c_char_p
        L.xmd_engine_create.argtypes = [C.POINTER(_abi.XmdParams), C.POINTER(_abi.XmdTopology), C.c_void_p]
        L.xmd_engine_destroy.argtypes = [C.c_void_p]
        L.xmd_step.argtypes = [C.c_void_p, C.c_int64]
        L.xmd_stream.restype = C.c_void_p
        L.xmd_stream.argtypes = [C.c_void_p]
        L.xmd_rand_normals.argtypes = [C.c_void_p, C.c_uint64, C.c_int64, C.POINTER(C.c_float),
                                       C.POINTER(C.c_uint64)]
        _LIB = L
    return _LIB


class Engine(HostSystem):
    """One engine over all replicas of `system` (an xmd_amd.synth.System)."""

    def __init__(self, params, system):
        super().__init__(params, system)
        L = lib()
        h = C.c_void_p()
        rc = L.xmd_engine_create(C.byref(params), C.byref(self.topo), C.byref(h))
        if rc != 0:
            raise XmdError(rc, L.xmd_last_error(None).decode())
        self.handle = h

    def _fn(self, name):
        return getattr(lib(), "xmd_" + name)

    def _err(self):
        return lib().xmd_last_error(self.handle).decode()

    def step(self, nsteps):
        self._check(lib().xmd_step(self.handle, C.c_int64(nsteps)))

    def stream(self):
        return lib().xmd_stream(self.handle)

    def replica_status(self, rep):
        """(status, message) of one replica: status 0, or the error that
        halted it (it stays halted until its state is set)."""
        st = lib().xmd_replica_status(self.handle, C.c_int32(rep))
        return st, (self._err() if st else "")

    def autotune(self, reps=10):
        """Pick the force kernel's lanes per bead (4, 8, 16) by timing;
        returns (chosen, [ms per step for 4, 8, 16])."""
        lpb = C.c_int32()
        ms = (C.c_double * 3)()
        self._check(lib().xmd_autotune(self.handle, C.c_int32(reps), C.byref(lpb), ms))
        return lpb.value, list(ms)

    def observables(self, active_threshold=1.5):
        """Per-replica stats-report observables computed on the GPU, one row
        per replica: t, V, K, E, Rg, W, active native contacts, native
        contacts (xmd/src/io/report_stats.cpp:8-41,
        xmd/src/io/report_structure.cpp:81-85)."""
        out = np.zeros(8 * self.nrep)
        self._check(lib().xmd_observables(self.handle, C.c_double(active_threshold),
                                          out.ctypes.data_as(C.POINTER(C.c_double))))
        return out.reshape(self.nrep, 8)

    def observables_noalloc(self, active_threshold=1.5):
        """observables() through the entry point that allocates no device
        memory (a fixed staging array of the library, at most 4096
        replicas)."""
        out = np.zeros(8 * self.nrep)
        self._check(lib().xmd_observables_noalloc(self.handle, C.c_double(active_threshold),
                                                  out.ctypes.data_as(C.POINTER(C.c_double))))
        return out.reshape(self.nrep, 8)

    def rand_normals(self, state, n):
        out = np.zeros(3 * max(n, 1), dtype=np.float32)
        so = C.c_uint64()
        self._check(lib().xmd_rand_normals(self.handle, C.c_uint64(state), C.c_int64(n),
                                           out.ctypes.data_as(C.POINTER(C.c_float)), C.byref(so)))
        return out[:3 * n].reshape(n, 3), so.value

    def close(self):
        if self.handle:
            lib().xmd_engine_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def profile(engine, nsteps):
    """Per-kernel HIP-event timing of nsteps eager steps:
    {name: (launches, total_ms)}."""
    cap = 64
    names = C.create_string_buffer(8192)
    counts = np.zeros(cap, dtype=np.int64)
    ms = np.zeros(cap)
    nk = C.c_int32()
    engine._check(lib().xmd_profile(engine.handle, C.c_int64(nsteps), names, C.c_int64(8192),
                                    counts.ctypes.data_as(C.POINTER(C.c_int64)),
                                    ms.ctypes.data_as(C.POINTER(C.c_double)), C.c_int32(cap), C.byref(nk)))
    lst = names.value.decode().split("\n")
    return {lst[k]: (int(counts[k]), float(ms[k])) for k in range(nk.value)}


def first_zero_log_draw(state, n_beads, n_steps):
    """(step, bead) of the first zero uniform draw that feeds log() in
    rand_gen's Langevin noise within n_steps, or None."""
    L = lib()
    L.xmd_rng_first_zero_draw.restype = C.c_int64
    L.xmd_rng_first_zero_draw.argtypes = [C.c_uint64, C.c_int64, C.c_uint32]
    idx = L.xmd_rng_first_zero_draw(C.c_uint64(int(state)), C.c_int64(4 * n_beads * n_steps), C.c_uint32(5))
    if idx < 0:
        return None
    return idx // (4 * n_beads), (idx // 4) % n_beads
