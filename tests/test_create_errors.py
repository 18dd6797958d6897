"""Argument checks of xmd_engine_create: every malformed xmd_params /
xmd_topology is refused with XMD_ERR_ARG and a message naming the problem,
before any device call, so these run without a GPU (the well-formed input
then fails only for the missing device, or creates an engine on a GPU).
The reference reports the same kinds of input by throwing from its loaders
(xmd/include/xmd/store/store.h:47, xmd/src/types/amino_acid.cpp:17,29)."""
import ctypes as C
import glob

import numpy as np
import pytest

from xmd_amd import _abi, synth
from xmd_amd.engine import lib


def base():
    """Two replicas, each of two 6-bead chains, with natives, a disulfide
    bond, a wall and AFM tips."""
    rng = np.random.default_rng(4)
    chains = [synth.random_chain(6, rng) for _ in range(2)]
    placed = synth.place_chains([c[0] for c in chains], 60.0, rng)
    rep = dict(chains=[(p, c[1]) for p, c in zip(placed, chains)], box_A=(60, 60, 60), temperature=0.35, seed=9)
    import dataclasses
    s = synth.make_system([rep, dict(rep, seed=10)])
    return dataclasses.replace(
        s, ss_i1=np.array([0, 12], dtype=np.int32), ss_i2=np.array([7, 19], dtype=np.int32),
        solid_walls=np.array([[0, 0, -5.0, 0, 0, 1]]), vel_afm_idx=np.array([3], dtype=np.int32),
        vel_afm_orig=np.full((1, 3), np.nan), vel_afm_v=np.zeros((1, 3)),
        force_afm_idx=np.array([20], dtype=np.int32), force_afm_F=np.ones((1, 3)),
        heur_bonded=np.array(([1, 3, 1, 0, 0, 0] * 4), dtype=np.uint8))


def create(params, topo):
    eng = C.c_void_p()
    st = lib().xmd_engine_create(C.byref(params), C.byref(topo), C.byref(eng))
    if st == _abi.OK:
        lib().xmd_engine_destroy(eng)
    return st, lib().xmd_last_error(None).decode()


def arr(t, name, values, ctype):
    a = np.ascontiguousarray(values, dtype=np.dtype(ctype))
    t._arrays[name] = a
    setattr(t, name, _abi.ptr(a, ctype))


def _int(t, name, values):
    arr(t, name, values, C.c_int32)


def _dbl(t, name, values):
    arr(t, name, values, C.c_double)


CASES = {
    # name: (mutation of (params, topology), expected message)
    "dt": (lambda p, t: setattr(p, "dt", 0.0), "dt must be positive"),
    "precision": (lambda p, t: setattr(p, "precision", 7), "bad precision"),
    "lanes": (lambda p, t: setattr(p, "force_lanes", 5), "force_lanes must be 0, 4, 8 or 16"),
    "groups": (lambda p, t: setattr(p, "stream_groups", -1), "stream_groups must be >= 0"),
    "row_width": (lambda p, t: setattr(p, "row_width", -5), "row_width must be >= 0"),
    "fp_capacity": (lambda p, t: setattr(p, "fp_capacity", -1), "fp_capacity must be >= 0"),
    "ct_capacity": (lambda p, t: setattr(p, "ct_capacity", -1), "ct_capacity must be >= 0"),
    "no beads": (lambda p, t: setattr(t, "n_beads", 0), "empty topology"),
    "null atype": (lambda p, t: setattr(t, "atype", C.cast(None, C.POINTER(C.c_int8))), "null topology array"),
    "replica span": (lambda p, t: _int(t, "replica_offset", [0, 12, 23]), "replica_offset must span"),
    "empty replica": (lambda p, t: (_int(t, "replica_offset", [0, 24, 24, 24]), setattr(t, "n_replicas", 3)),
                      "empty replica"),
    "chain span": (lambda p, t: _int(t, "chain_offset", [0, 6, 12, 18, 23]), "chain_offset must span"),
    "empty chain": (lambda p, t: _int(t, "chain_offset", [0, 6, 6, 18, 24]), "empty chain"),
    "chain across replicas": (lambda p, t: _int(t, "chain_offset", [0, 6, 14, 18, 24]), "chain spans replicas"),
    "atype": (lambda p, t: arr(t, "atype", [20] + [0] * 23, C.c_int8), "atype out of range"),
    "nan position": (lambda p, t: t._arrays["r0"].__setitem__(5, np.nan), "non-finite initial position"),
    "box": (lambda p, t: t._arrays["box"].__setitem__(4, -1.0), "box lengths must be finite and >= 0"),
    "native order": (lambda p, t: (_int(t, "nat_i1", [5]), _int(t, "nat_i2", [2]), _dbl(t, "nat_dist", [1.0]),
                                   setattr(t, "n_native", 1)), "native contact needs 0 <= i1 < i2 < n"),
    "native bounds": (lambda p, t: (_int(t, "nat_i1", [-1]), _int(t, "nat_i2", [2]), _dbl(t, "nat_dist", [1.0]),
                                    setattr(t, "n_native", 1)), "native contact needs 0 <= i1 < i2 < n"),
    "native across replicas": (lambda p, t: (_int(t, "nat_i1", [2]), _int(t, "nat_i2", [15]),
                                             _dbl(t, "nat_dist", [1.0]), setattr(t, "n_native", 1)),
                               "native contact crosses replicas"),
    "native distance": (lambda p, t: (_int(t, "nat_i1", [0]), _int(t, "nat_i2", [4]), _dbl(t, "nat_dist", [0.0]),
                                      setattr(t, "n_native", 1)), "native contact distance must be positive"),
    "null natives": (lambda p, t: (setattr(t, "n_native", 3), setattr(t, "nat_dist", C.cast(None, C.POINTER(C.c_double)))),
                     "bad native contact arrays"),
    "ssbond across replicas": (lambda p, t: _int(t, "ss_i2", [13, 19]), "ssbond crosses replicas"),
    "ssbond order": (lambda p, t: _int(t, "ss_i1", [8, 12]), "ssbond needs 0 <= i1 < i2 < n"),
    "walls": (lambda p, t: (setattr(t, "n_lj_walls", 2), setattr(t, "lj_walls", C.cast(None, C.POINTER(C.c_double)))),
              "bad wall arrays"),
    "afm arrays": (lambda p, t: setattr(t, "n_force_afm", -1), "bad AFM arrays"),
    "afm index": (lambda p, t: _int(t, "vel_afm_idx", [24]), "velocity AFM tip outside the beads"),
    "heuristic angle": (lambda p, t: arr(t, "heur_bonded", [0, 0, 0, 0, 1] + [0] * 19, C.c_uint8),
                        "heuristic angle does not join consecutive residues"),
    "heuristic dihedral": (lambda p, t: arr(t, "heur_bonded", [0, 0, 0, 2] + [0] * 20, C.c_uint8),
                           "heuristic dihedral does not join consecutive residues"),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_malformed_input_is_refused(case):
    system = base()
    params = synth.config_params(2)
    topo = system.topology()
    mutate, msg = CASES[case]
    mutate(params, topo)
    st, err = create(params, topo)
    assert st == _abi.ERR_ARG, (st, err)
    assert msg in err


def test_well_formed_input_passes_the_checks():
    system = base()
    st, err = create(synth.config_params(2), system.topology())
    if glob.glob("/dev/kfd"):
        assert st == _abi.OK, err
    else:
        assert st == _abi.ERR_NO_DEVICE, err
