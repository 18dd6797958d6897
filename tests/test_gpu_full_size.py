"""Parity at the workloads' full sizes (FP64 engine against the oracle):

* C1 (1000 beads, open box), the per-term protocol: forces and energies of
  every term on 100 seeded conformations (the native chain plus isotropic
  0.3 A Gaussian displacements, seeds 0..99), lists bit-exact;
* C2 (40 x 250 beads, periodic, full term set): every list bit-exact,
  per-term forces, 10 steps of trajectory with QA state;
* C3 (512 single chains of 30..400 beads): every replica's lists and
  per-term forces, 5 steps of the batch against the oracle's replicas;
* C4 (400 x 250 beads, periodic): every list bit-exact and per-term forces
  on one conformation (the oracle's O(N^2) list build is the cost).
Bars as in test_gpu_parity: lists equal in order, forces within 1e-10
relative, positions within 1e-8.
"""
import numpy as np
import pytest

from xmd_amd import synth, _abi
from xmd_amd.engine import Engine
from oracle.oracle import Oracle
from test_gpu_parity import compare_terms, lists_equal

pytestmark = pytest.mark.gpu


def _displaced(system, seed, rep=None):
    sl = slice(None) if rep is None else system.replica_slice(rep)
    rng = np.random.default_rng(seed)
    r = system.r0[:, sl]
    return r + rng.normal(scale=0.3 * synth.ANGSTROM, size=r.shape)


def test_c1_hundred_conformations():
    system = synth.config_system(1)
    params = synth.config_params(1)
    print("C1", system.n_beads, flush=True)
    eng, ora = Engine(params, system), Oracle(params, system)
    for seed in range(100):
        r = _displaced(system, seed)
        eng.set_positions(0, r)
        ora.set_positions(0, r)
        eng.rebuild_lists()
        ora.rebuild_lists()
        if seed % 10 == 0:
            lists_equal(eng, ora, 0)
        compare_terms(eng, ora)
    eng.close()


def test_c2_full_box():
    system = synth.config_system(2)
    params = synth.config_params(2)
    print("C2", system.n_beads, flush=True)
    eng, ora = Engine(params, system), Oracle(params, system)
    r = _displaced(system, 1)
    eng.set_positions(0, r)
    ora.set_positions(0, r)
    eng.rebuild_lists()
    ora.rebuild_lists()
    lists_equal(eng, ora, 0)
    compare_terms(eng, ora)
    eng.step(10)
    ora.step(10)
    y, _, rng, t = eng.get_state(0)
    yo, _, rngo, to = ora.get_state(0)
    assert rng == rngo and t == to
    assert np.abs(y - yo).max() < 1e-8
    ce, se = eng.get_qa_state(0)
    co, so = ora.get_qa_state(0)
    assert len(ce) == len(co) and np.array_equal(se, so)
    eng.close()


def test_c3_all_replicas():
    system = synth.config_system(3)
    params = synth.config_params(3)
    print("C3", system.n_replicas, system.n_beads, flush=True)
    eng, ora = Engine(params, system), Oracle(params, system)
    for k in range(system.n_replicas):  # off the native geometry (tethers at rest)
        r = _displaced(system, 1000 + k, rep=k)
        eng.set_positions(k, r)
        ora.set_positions(k, r)
    eng.rebuild_lists()
    ora.rebuild_lists()
    for k in range(system.n_replicas):
        lists_equal(eng, ora, k, names=("nl_all", "go", "pauli", "es"))
    compare_terms(eng, ora, rep_count=system.n_replicas)
    eng.step(5)
    ora.step(5)
    for k in range(system.n_replicas):
        y, _, rng, t = eng.get_state(k)
        yo, _, rngo, to = ora.get_state(k)
        assert rng == rngo and t == to
        assert np.abs(y - yo).max() < 1e-8, k
    eng.close()


def test_c4_full_box():
    system = synth.config_system(4)
    params = synth.config_params(4)
    print("C4", system.n_beads, flush=True)
    eng, ora = Engine(params, system), Oracle(params, system)
    r = _displaced(system, 2)
    eng.set_positions(0, r)
    ora.set_positions(0, r)
    eng.rebuild_lists()
    ora.rebuild_lists()
    lists_equal(eng, ora, 0)
    compare_terms(eng, ora)
    eng.close()
