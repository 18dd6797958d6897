"""Parity at the workloads' full sizes (FP64 engine against the oracle):

* C1 (1000 beads, open box), the per-term protocol: forces and energies of
  every term on 100 seeded conformations (the native chain plus isotropic
  0.3 A Gaussian displacements, seeds 0..99), lists bit-exact;
* C2 (40 x 250 beads, periodic, full term set): every list bit-exact,
  per-term forces, 10 steps of trajectory with QA state;
* C3 (512 single chains of 30..400 beads): every replica's lists and
  per-term forces, 5 steps of the batch against the oracle's replicas;
* C4 (400 x 250 beads, periodic): every list bit-exact and per-term forces
  on one conformation, then 5 steps of trajectory with QA state (the
  oracle's O(N^2) list build is the cost).
* C1 in mixed precision (FP32 pair and bonded terms): every term's forces
  on the 100 conformations within 1e-5 relative RMS of the FP64 oracle, and
  its energy within 1e-5 of the term's scale;
* C4's batch of boxes at T = 0.30 + 0.2 k / 7 (8 replicas of one box at
  C4's density and term set, 2,000 beads each so the oracle can run them),
  each replica against the oracle running that box alone.
Bars as in test_gpu_parity: lists equal in order, forces within 1e-10
relative, positions within 1e-8, QA contacts equal field by field.
"""
import numpy as np
import pytest

from xmd_amd import synth, _abi
from xmd_amd.engine import Engine
from oracle.oracle import Oracle
from test_gpu_parity import compare_terms, contacts_equal, lists_equal

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
    assert contacts_equal(eng, ora) > 0
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
    print("C4 steps", flush=True)
    eng.step(5)
    ora.step(5)
    y, _, rng, t = eng.get_state(0)
    yo, _, rngo, to = ora.get_state(0)
    assert rng == rngo and t == to
    assert np.abs(y - yo).max() < 1e-8
    assert contacts_equal(eng, ora) > 0
    eng.close()


def test_c1_mixed_hundred_conformations():
    """C1's FP32-force protocol: per-term forces and energies of the mixed
    engine against the FP64 oracle on the 100 seeded conformations."""
    from test_gpu_parity import rel_rms
    system = synth.config_system(1)
    eng = Engine(synth.config_params(1, precision=_abi.PREC_MIXED), system)
    ora = Oracle(synth.config_params(1), system)
    terms = ["tether", "chiral", "nat_ang", "cnd", "pauli", "go", "const_es"]  # C1's term set
    worst_F = {t: 0.0 for t in terms}
    worst_V = {t: 0.0 for t in terms}
    for seed in range(100):
        r = _displaced(system, seed)
        eng.set_positions(0, r)
        ora.set_positions(0, r)
        for t in terms:
            m = 1 << _abi.TERM[t]
            Fo, Vo = ora.eval_forces(m)
            if np.abs(Fo).max() == 0:
                continue
            Fe, Ve = eng.eval_forces(m)
            k = _abi.TERM[t]
            worst_F[t] = max(worst_F[t], rel_rms(Fe, Fo))
            # the term's scale: its energy's magnitude, at least its mean
            # force times a bead's displacement scale over all beads
            scale = max(abs(Vo[0, k]), 1e-3 * np.sqrt((Fo ** 2).sum()))
            worst_V[t] = max(worst_V[t], abs(Ve[0, k] - Vo[0, k]) / scale)
    print("C1 mixed vs FP64 oracle, worst over 100 conformations: F", worst_F, "V", worst_V, flush=True)
    for t in terms:
        assert worst_F[t] < 1e-5, (t, worst_F[t])
        assert worst_V[t] < 1e-5, (t, worst_V[t])
    eng.close()


def test_c4_temperature_batch():
    """C4's batched mode: boxes of C4's density and full term set at
    T = 0.30 + 0.2 k / 7 with their own seeds, stepped as one batch; each
    replica equals the oracle running its box alone."""
    base = synth.config_system(4, chains=8, chain_len=250)
    temps = [0.30 + 0.2 * k / 7 for k in range(8)]
    seeds = [synth.SEED_BASE + 4 + 1 + k for k in range(8)]
    batch = synth.replicate(base, temps, seeds)
    params = synth.config_params(4)
    print("C4 batch", batch.n_replicas, batch.n_beads, flush=True)
    eng = Engine(params, batch)
    eng.step(30)
    for k in range(batch.n_replicas):
        one = synth.replicate(base, [temps[k]], [seeds[k]])
        ora = Oracle(params, one)
        ora.step(30)
        y, _, rng, t = eng.get_state(k)
        yo, _, rngo, to = ora.get_state(0)
        assert rng == rngo and t == to, k
        assert np.abs(y - yo).max() < 1e-8, (k, np.abs(y - yo).max())
        contacts_equal(eng, ora, rep=k, ora_rep=0)
        ora.close()
    eng.close()


@pytest.mark.parametrize("cid", [2, 3, 4])
def test_mixed_precision_full_size(cid):
    """Mixed-precision forces within the 1e-5 rel-RMS bar of the FP64
    engine at the full sizes, per term and in total (C2, all 512 C3
    replicas, the 100k-bead C4 box)."""
    from test_gpu_parity import rel_rms
    system = synth.config_system(cid)
    print("mixed", cid, system.n_beads, flush=True)
    e64 = Engine(synth.config_params(cid), system)
    emx = Engine(synth.config_params(cid, precision=_abi.PREC_MIXED), system)
    for k in range(system.n_replicas):
        r = _displaced(system, 3000 + k, rep=k)
        e64.set_positions(k, r)
        emx.set_positions(k, r)
    for name in _abi.TERMS:
        m = 1 << _abi.TERM[name]
        F64, _ = e64.eval_forces(m)
        if np.abs(F64).max() == 0:
            continue
        Fmx, _ = emx.eval_forces(m)
        assert rel_rms(Fmx, F64) < 1e-5, (name, rel_rms(Fmx, F64))
    F64, V64 = e64.eval_forces()
    Fmx, Vmx = emx.eval_forces()
    assert rel_rms(Fmx, F64) < 1e-5
    assert np.allclose(Vmx.sum(axis=1), V64.sum(axis=1), rtol=1e-5, atol=1e-6)
    e64.close()
    emx.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_c1_runs_its_step_count(prec):
    """C1 (one 1000-bead chain, open box) stepped for its stated 100,000
    steps (a seed whose noise has no zero draw in that horizon): every
    replica step is taken, the state stays finite, the chain keeps its
    bonds (every tether within 20 % of its native length) and the kinetic
    energy per bead stays at the bath's scale -- 3/2 <v^2> with the
    stationary <v^2> = 1.22 T dt per component of the reference's noise,
    which lacks the dt^-1/2 factor (lang_pc.cpp:15-19)."""
    system = synth.config_system(1, horizon_steps=100000)
    p = synth.config_params(1, precision={"fp64": _abi.PREC_FP64, "mixed": _abi.PREC_MIXED}[prec])
    eng = Engine(p, system)
    ks = []
    for _ in range(10):
        eng.step(10000)
        obs = eng.observables()[0]  # t, V, K, E, Rg, W, active natives, natives
        ks.append(obs[2] / system.n_beads)
    c = eng.counters()
    y, _, _, t = eng.get_state(0)
    print(prec, "t", t, "K per bead", [round(k, 4) for k in ks], "rebuilds", c["rebuilds"], flush=True)
    assert c["steps"] == 100000 and eng.replica_status(0)[0] == 0
    assert np.isfinite(y).all()
    r = y[0]
    bond = np.linalg.norm(r[:, 1:] - r[:, :-1], axis=0)
    assert np.abs(bond / (synth.BOND * synth.ANGSTROM) - 1).max() < 0.2
    k_bath = 1.5 * 1.22 * float(system.temperature[0]) * p.dt
    assert 0.3 * k_bath < np.mean(ks) < 3.0 * k_bath, (np.mean(ks), k_bath)
    eng.close()
