"""Heuristic angle and dihedral terms on the GPU: the glut sequence-file
model (2 replicas; 6356 of its angles and 6362 of its dihedrals have no
native value) with every default term plus both heuristic terms, against
the oracle.  Bars as the other parity tests: FP64 forces per term within
1e-10 relative, trajectories step for step within 1e-8; mixed-precision
rel-RMS within 1e-5 of the FP64 engine per heuristic term and in total."""
import os

import numpy as np
import pytest

from xmd_amd import _abi, model as M, params as P
from xmd_amd.engine import Engine
from oracle.oracle import Oracle
from test_gpu_parity import compare_terms, lists_equal, rel_rms

pytestmark = pytest.mark.gpu

MODELS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models")
HEUR = {"heurestic angles": {"enabled": True}, "heurestic dihedrals": {"enabled": True}}


def glut(precision=_abi.PREC_FP64):
    params, system, _ = M.from_input(dict(HEUR, model={"seqfile": "glut/glut.yml"}), base_dir=MODELS, replicas=2)
    params.precision = precision
    assert params.heur_ang_enabled and params.heur_dih_enabled
    return params, system


def displaced(system, seed=7):
    rng = np.random.default_rng(seed)
    return system.r0 + rng.normal(scale=0.3 * P.ANGSTROM, size=system.r0.shape)


def test_heuristic_forces_fp64():
    params, system = glut()
    print("glut heuristic fp64", system.n_beads, flush=True)
    eng, ora = Engine(params, system), Oracle(params, system)
    r = displaced(system)
    for k in range(system.n_replicas):
        sl = system.replica_slice(k)
        eng.set_positions(k, r[:, sl])
        ora.set_positions(k, r[:, sl])
    eng.rebuild_lists()
    ora.rebuild_lists()
    lists_equal(eng, ora, 0)
    for name in ("heur_ang", "heur_dih"):
        F, V = ora.eval_forces(1 << _abi.TERM[name])
        assert np.abs(F).max() > 0 and np.all(V[:, _abi.TERM[name]] != 0), name
    compare_terms(eng, ora, rep_count=system.n_replicas)
    eng.close()


def test_heuristic_trajectory_fp64():
    params, system = glut()
    print("glut heuristic trajectory", system.n_beads, flush=True)
    eng, ora = Engine(params, system), Oracle(params, system)
    eng.step(20)
    ora.step(20)
    for k in range(system.n_replicas):
        y, _, rng, t = eng.get_state(k)
        yo, _, rngo, to = ora.get_state(k)
        assert rng == rngo and t == to
        assert np.abs(y - yo).max() < 1e-8
    eng.close()


def test_heuristic_mixed_precision():
    params, system = glut()
    pm, _ = glut(_abi.PREC_MIXED)
    print("glut heuristic mixed", system.n_beads, flush=True)
    e64, emx = Engine(params, system), Engine(pm, system)
    r = displaced(system, 11)
    for k in range(system.n_replicas):
        sl = system.replica_slice(k)
        e64.set_positions(k, r[:, sl])
        emx.set_positions(k, r[:, sl])
    for name in ("heur_ang", "heur_dih"):
        m = 1 << _abi.TERM[name]
        F64, V64 = e64.eval_forces(m)
        Fmx, Vmx = emx.eval_forces(m)
        assert rel_rms(Fmx, F64) < 1e-5, (name, rel_rms(Fmx, F64))
        assert np.allclose(Vmx[:, _abi.TERM[name]], V64[:, _abi.TERM[name]], rtol=1e-5), name
    F64, _ = e64.eval_forces()
    Fmx, _ = emx.eval_forces()
    assert rel_rms(Fmx, F64) < 1e-5
    e64.close()
    emx.close()
