"""The engine on models loaded from the reference's example inputs, under
the default parameters (every step-path term the defaults enable): the
glut sequence file morphed into random chains (2 replicas, seeds 42, 43)
and the 1ubq PDB structure in its CRYST1 cell.  Same bars as the other
parity tests: lists bit-exact, FP64 forces per term within 1e-10 of the
oracle, FP64 trajectories step for step, mixed-precision force rel-RMS
within 1e-5."""
import os

import numpy as np
import pytest

from xmd_amd import _abi, model as M, params as P
from xmd_amd.engine import Engine
from oracle.oracle import Oracle
from test_gpu_parity import compare_terms, lists_equal, rel_rms

pytestmark = pytest.mark.gpu

MODELS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models")


@pytest.fixture(scope="module")
def glut():
    params, system, _ = M.from_input({"model": {"seqfile": "glut/glut.yml"}}, base_dir=MODELS, replicas=2)
    return params, system


@pytest.fixture(scope="module")
def ubq():
    tree, _ = P.load(None)
    del tree["model"]["seqfile"], tree["model"]["morph into SAW"]
    tree["model"]["pdb"] = "1ubq/1ubq.pdb"
    m, state = M.load_model(tree, MODELS)
    params = P.to_xmd_params(tree, MODELS)
    return params, M.to_system(m, P.quantity(tree["langevin"]["temperature"]), state)


@pytest.mark.parametrize("case", ["glut", "ubq"])
def test_model_lists_and_forces_fp64(case, glut, ubq):
    params, system = glut if case == "glut" else ubq
    print("case", case, system.n_beads, flush=True)
    eng, ora = Engine(params, system), Oracle(params, system)
    # displaced by 0.3 A (Gaussian): the morphed bonds sit exactly at the
    # default tether length, where the tether force is round-off
    rng = np.random.default_rng(7)
    for r in range(system.n_replicas):
        sl = system.replica_slice(r)
        pos = system.r0[:, sl] + rng.normal(scale=0.3 * P.ANGSTROM, size=system.r0[:, sl].shape)
        eng.set_positions(r, pos)
        ora.set_positions(r, pos)
    eng.rebuild_lists()
    ora.rebuild_lists()
    for r in range(system.n_replicas):
        lists_equal(eng, ora, r)
    compare_terms(eng, ora, rep_count=system.n_replicas)
    eng.close()


@pytest.mark.parametrize("case", ["glut", "ubq"])
def test_model_trajectory_fp64(case, glut, ubq):
    params, system = glut if case == "glut" else ubq
    print("case", case, system.n_beads, flush=True)
    eng, ora = Engine(params, system), Oracle(params, system)
    eng.step(20)
    ora.step(20)
    for r in range(system.n_replicas):
        y, _, rng, t = eng.get_state(r)
        yo, _, rngo, to = ora.get_state(r)
        assert rng == rngo and t == to
        assert np.abs(y - yo).max() < 1e-8
        ce, se = eng.get_qa_state(r)
        co, so = ora.get_qa_state(r)
        assert len(ce) == len(co) and np.array_equal(se, so)
    eng.close()


def test_model_mixed_precision(glut):
    params, system = glut
    params.precision = _abi.PREC_MIXED
    eng = Engine(params, system)
    params.precision = _abi.PREC_FP64
    ref = Engine(params, system)
    eng.step(50)
    # the same state in both: Gear derivatives, generator, time, QA contacts
    for r in range(system.n_replicas):
        y, _, rng, t = eng.get_state(r)
        cont, sync = eng.get_qa_state(r)
        for e in (ref, eng):
            e.set_state(r, y, rng, t)
            e.set_qa_state(r, cont, sync)
    F, _ = eng.eval_forces()
    Fr, _ = ref.eval_forces()
    assert rel_rms(F, Fr) < 1e-5, rel_rms(F, Fr)
    eng.close()
    ref.close()
