"""Host-side wrapper over a C library that implements the xmd_hip.h
operations (the engine's `xmd_*` entry points).  Returns numpy arrays;
raises XmdError on a non-zero status."""
import ctypes as C

import numpy as np

from . import _abi


class XmdError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"xmd status {status}: {msg}")
        self.status = status


CONTACT_DTYPE = np.dtype([("i1", np.int32), ("i2", np.int32), ("type", np.int16), ("status", np.int8),
                          ("_pad", np.int8), ("diff1", np.int8, 4), ("diff2", np.int8, 4),
                          ("ref_time", np.float64)], align=True)
assert CONTACT_DTYPE.itemsize == C.sizeof(_abi.XmdQaContact)


class HostSystem:
    """Common driver: subclasses bind `_fn(name)` to the library symbols."""

    def __init__(self, params, system):
        self.params = params
        self.system = system
        self.topo = system.topology()
        self.n = system.n_beads
        self.nrep = system.n_replicas
        self.handle = None

    # subclass hooks
    def _fn(self, name):
        raise NotImplementedError

    def _err(self):
        raise NotImplementedError

    def _check(self, rc):
        if rc != 0:
            raise XmdError(rc, self._err())

    def _rn(self, rep):
        sl = self.system.replica_slice(rep)
        return sl.stop - sl.start

    def set_state(self, rep, y, rng_state, t):
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1)
        assert y.size == 18 * self._rn(rep)
        self._check(self._fn("set_state")(self.handle, rep, _abi.ptr(y, C.c_double), C.c_uint64(int(rng_state)),
                                          C.c_double(t)))

    def get_state(self, rep):
        n = self._rn(rep)
        y = np.zeros(18 * n)
        v = np.zeros(3 * n)
        rng = C.c_uint64()
        t = C.c_double()
        self._check(self._fn("get_state")(self.handle, rep, _abi.ptr(y, C.c_double), _abi.ptr(v, C.c_double),
                                          C.byref(rng), C.byref(t)))
        return y.reshape(6, 3, n), v.reshape(3, n), rng.value, t.value

    def set_positions(self, rep, r):
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1)
        assert r.size == 3 * self._rn(rep)
        self._check(self._fn("set_positions")(self.handle, rep, _abi.ptr(r, C.c_double)))

    def rebuild_lists(self):
        self._check(self._fn("rebuild_lists")(self.handle))

    def get_list(self, rep, name):
        lid = _abi.LIST[name] if isinstance(name, str) else int(name)
        n = C.c_int64()
        self._check(self._fn("get_list")(self.handle, rep, lid, None, None, C.c_int64(0), C.byref(n)))
        i1 = np.zeros(max(n.value, 1), dtype=np.int32)
        i2 = np.zeros(max(n.value, 1), dtype=np.int32)
        self._check(self._fn("get_list")(self.handle, rep, lid, _abi.ptr(i1, C.c_int32), _abi.ptr(i2, C.c_int32),
                                         C.c_int64(n.value), C.byref(n)))
        return i1[:n.value], i2[:n.value]

    def eval_forces(self, mask=_abi.ALL_TERMS):
        F = np.zeros(3 * self.n)
        V = np.zeros(self.nrep * _abi.NUM_TERMS)
        self._check(self._fn("eval_forces")(self.handle, C.c_uint32(mask), _abi.ptr(F, C.c_double),
                                            _abi.ptr(V, C.c_double)))
        return F.reshape(3, self.n), V.reshape(self.nrep, _abi.NUM_TERMS)

    def qa_update(self):
        self._check(self._fn("qa_update")(self.handle))

    def get_qa_state(self, rep):
        n = C.c_int64()
        nb = self._rn(rep)
        sync = np.zeros(4 * nb, dtype=np.int8)
        self._check(self._fn("get_qa_state")(self.handle, rep, None, C.c_int64(0), C.byref(n),
                                             _abi.ptr(sync, C.c_int8)))
        cont = np.zeros(max(n.value, 1), dtype=CONTACT_DTYPE)
        self._check(self._fn("get_qa_state")(self.handle, rep, cont.ctypes.data_as(C.POINTER(_abi.XmdQaContact)),
                                             C.c_int64(n.value), C.byref(n), _abi.ptr(sync, C.c_int8)))
        return cont[:n.value], sync.reshape(nb, 4)

    def set_qa_state(self, rep, contacts, sync):
        contacts = np.ascontiguousarray(contacts, dtype=CONTACT_DTYPE)
        sync = np.ascontiguousarray(sync, dtype=np.int8).reshape(-1)
        self._check(self._fn("set_qa_state")(self.handle, rep,
                                             contacts.ctypes.data_as(C.POINTER(_abi.XmdQaContact)),
                                             C.c_int64(len(contacts)), _abi.ptr(sync, C.c_int8)))

    def get_wall_state(self, rep):
        """LJ attractive wall state (status, joint, ref_t), per bead x wall."""
        k = self._rn(rep) * len(self.system.lj_walls)
        st = np.zeros(max(k, 1), dtype=np.int8)
        jo = np.zeros(max(3 * k, 1))
        rt = np.zeros(max(k, 1))
        self._check(self._fn("get_wall_state")(self.handle, rep, _abi.ptr(st, C.c_int8), _abi.ptr(jo, C.c_double),
                                               _abi.ptr(rt, C.c_double)))
        return st[:k], jo[:3 * k].reshape(k, 3), rt[:k]

    def set_wall_state(self, rep, status, joint, ref_t):
        st = np.ascontiguousarray(status, dtype=np.int8)
        jo = np.ascontiguousarray(joint, dtype=np.float64).reshape(-1)
        rt = np.ascontiguousarray(ref_t, dtype=np.float64)
        assert st.size == self._rn(rep) * len(self.system.lj_walls) and jo.size == 3 * st.size
        self._check(self._fn("set_wall_state")(self.handle, rep, _abi.ptr(st, C.c_int8), _abi.ptr(jo, C.c_double),
                                               _abi.ptr(rt, C.c_double)))

    def energy(self):
        V = np.zeros(self.nrep)
        self._check(self._fn("get_energy")(self.handle, _abi.ptr(V, C.c_double)))
        return V

    def counters(self):
        out = np.zeros(5, dtype=np.int64)
        self._check(self._fn("get_counters")(self.handle, _abi.ptr(out, C.c_int64)))
        return dict(steps=int(out[0]), rebuilds=int(out[1]), nl_entries=int(out[2]), nl_capacity=int(out[3]),
                    qa_contacts=int(out[4]))

    def positions(self, rep):
        y, _, _, _ = self.get_state(rep)
        return y[0]
