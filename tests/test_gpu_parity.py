"""Parity of the HIP engine (through the C ABI) with the CPU oracle.

Bars: pair lists, QA contact state and Langevin noise bit-exact on the same
inputs; FP64 forces and energies within 1e-10 relative (summation order and
device libm); mixed precision total force rel-RMS <= 1e-5 vs the FP64 oracle.
"""
import numpy as np
import pytest

from xmd_amd import synth, _abi
from xmd_amd.engine import Engine
from oracle.oracle import Oracle, rand_normals

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    den = np.sqrt((b ** 2).sum())
    return np.sqrt(((a - b) ** 2).sum()) / den if den > 0 else np.sqrt(((a - b) ** 2).sum())


def lists_equal(eng, ora, rep, names=_abi.LISTS):
    for name in names:
        a1, a2 = eng.get_list(rep, name)
        b1, b2 = ora.get_list(rep, name)
        assert len(a1) == len(b1), (name, len(a1), len(b1))
        assert np.array_equal(a1, b1) and np.array_equal(a2, b2), name


@pytest.fixture(scope="module")
def c1_small():
    system = synth.config_system(1, chain_len=300)
    params = synth.config_params(1)
    return params, system


def test_rand_normals_bit_exact(c1_small):
    params, system = c1_small
    eng = Engine(params, system)
    for state, n in [(0x5EED0001, 1000), (12345678901234567, 70001), (1, 64)]:
        g, gs = eng.rand_normals(state, n)
        o, os_ = rand_normals(state, n)
        assert gs == os_
        bad = np.flatnonzero(g.view(np.uint32) != o.view(np.uint32))
        assert bad.size == 0, (state, bad[:5], g.reshape(-1)[bad[:5]], o.reshape(-1)[bad[:5]])


def test_lists_bit_exact_c1(c1_small):
    params, system = c1_small
    eng, ora = Engine(params, system), Oracle(params, system)
    ora.rebuild_lists()
    lists_equal(eng, ora, 0)


def compare_terms(eng, ora, rep_count=1, tol=1e-10):
    for k, name in enumerate(_abi.TERMS):
        F, V = eng.eval_forces(1 << k)
        Fo, Vo = ora.eval_forces(1 << k)
        if np.abs(Fo).max() == 0 and np.abs(F).max() == 0:
            continue
        assert rel_rms(F, Fo) < tol, (name, rel_rms(F, Fo))
        for r in range(rep_count):
            assert abs(V[r, k] - Vo[r, k]) <= tol * max(1.0, abs(Vo[r, k])), name
    F, V = eng.eval_forces()
    Fo, Vo = ora.eval_forces()
    assert rel_rms(F, Fo) < tol


def test_forces_per_term_fp64_c1(c1_small):
    """C1 protocol: native structure + isotropic 0.3 A Gaussian displacement."""
    params, system = c1_small
    eng, ora = Engine(params, system), Oracle(params, system)
    for seed in range(3):
        r = synth.perturbed_positions(system, seed)
        eng.set_positions(0, r)
        ora.set_positions(0, r)
        compare_terms(eng, ora)


def test_trajectory_fp64_c1(c1_small):
    params, system = c1_small
    eng, ora = Engine(params, system), Oracle(params, system)
    for nsteps in (1, 1, 18):
        eng.step(nsteps)
        ora.step(nsteps)
        y, _, rng, t = eng.get_state(0)
        yo, _, rngo, to = ora.get_state(0)
        assert rng == rngo
        assert t == to
        assert np.abs(y - yo).max() < 1e-9
    assert eng.counters()["rebuilds"] == ora.counters()["rebuilds"]
    assert abs(eng.energy()[0] - ora.energy()[0]) < 1e-9 * max(1, abs(ora.energy()[0]))


@pytest.fixture(scope="module")
def c2_small():
    system = synth.config_system(2, chains=6, chain_len=80)
    params = synth.config_params(2)
    return params, system


def test_lists_bit_exact_pbc(c2_small):
    params, system = c2_small
    eng, ora = Engine(params, system), Oracle(params, system)
    ora.rebuild_lists()
    lists_equal(eng, ora, 0)


def test_forces_per_term_fp64_pbc(c2_small):
    params, system = c2_small
    eng, ora = Engine(params, system), Oracle(params, system)
    # a stretch of dynamics on the oracle gives QA contacts and PID activity
    ora.step(300)
    y, _, rng, t = ora.get_state(0)
    cont, sync = ora.get_qa_state(0)
    eng.set_state(0, y, rng, t)
    eng.set_qa_state(0, cont, sync)
    ora.set_state(0, y, rng, t)
    r = y[0] + np.random.default_rng(5).normal(scale=0.3 * 0.2, size=y[0].shape)
    eng.set_positions(0, r)
    ora.set_positions(0, r)
    compare_terms(eng, ora)


def test_qa_state_bit_exact_pbc(c2_small):
    params, system = c2_small
    eng, ora = Engine(params, system), Oracle(params, system)
    # drive the oracle; after each step copy its exact state into the engine
    # and compare one QA bookkeeping pass on identical coordinates
    for it in range(6):
        y, _, rng, t = ora.get_state(0)
        cont, sync = ora.get_qa_state(0)
        eng.set_state(0, y, rng, t)
        ora.set_state(0, y, rng, t)
        eng.set_qa_state(0, cont, sync)
        ora.set_qa_state(0, cont, sync)
        eng.qa_update()
        ora.qa_update()
        ce, se = eng.get_qa_state(0)
        co, so = ora.get_qa_state(0)
        assert np.array_equal(se, so)
        assert len(ce) == len(co)
        assert ce.tobytes() == co.tobytes()
        lists_equal(eng, ora, 0, ["qa_free"])
        ora.step(200)


def contacts_equal(eng, ora, rep=0, ora_rep=None):
    """QA contact sets (i1, i2, type, status, reference time, budget
    differences, in set order) and budgets of replica `rep` of the engine
    and `ora_rep` (default the same) of the oracle equal; returns the
    number of contacts."""
    ce, se = eng.get_qa_state(rep)
    co, so = ora.get_qa_state(rep if ora_rep is None else ora_rep)
    assert np.array_equal(se, so), rep
    assert len(ce) == len(co), (rep, len(ce), len(co))
    for f in ("i1", "i2", "type", "status", "diff1", "diff2", "ref_time"):
        assert np.array_equal(ce[f], co[f]), (rep, f)
    return len(ce)


def test_trajectory_fp64_pbc(c2_small):
    params, system = c2_small
    eng, ora = Engine(params, system), Oracle(params, system)
    eng.step(10)
    ora.step(10)
    y, _, rng, t = eng.get_state(0)
    yo, _, rngo, to = ora.get_state(0)
    assert rng == rngo and t == to
    assert np.abs(y - yo).max() < 1e-8
    contacts_equal(eng, ora)


def test_mixed_precision_force_error(c2_small):
    params, system = c2_small
    pm = synth.config_params(2, precision=_abi.PREC_MIXED)
    eng = Engine(pm, system)
    ora = Oracle(params, system)
    r = synth.perturbed_positions(system, 1)
    eng.set_positions(0, r)
    ora.set_positions(0, r)
    F, _ = eng.eval_forces()
    Fo, _ = ora.eval_forces()
    assert rel_rms(F, Fo) < 1e-5


def test_batched_replicas_match_single():
    system = synth.config_system(3, n_replicas=6)
    params = synth.config_params(3)
    eng = Engine(params, system)
    eng.step(12)
    ora = Oracle(params, system)
    ora.step(12)
    for k in range(system.n_replicas):
        y, _, rng, t = eng.get_state(k)
        yo, _, rngo, to = ora.get_state(k)
        assert rng == rngo and t == to
        assert np.abs(y - yo).max() < 1e-9


@pytest.mark.parametrize("prec", [_abi.PREC_FP64, _abi.PREC_MIXED])
def test_stream_groups_do_not_change_results(prec):
    """Replicas stepped as one part or as two concurrent stream groups give
    bit-identical forces, energies, lists, QA state and trajectories: no
    kernel mixes replicas."""
    system = synth.config_system(2, n_batch=3, chains=6, horizon_steps=80)
    engs = []
    for groups in (1, 2):
        params = synth.config_params(2, precision=prec)
        params.stream_groups = groups
        engs.append(Engine(params, system))
    a, b = engs
    Fa, Va = a.eval_forces()
    Fb, Vb = b.eval_forces()
    assert np.array_equal(Fa, Fb) and np.array_equal(Va, Vb)
    a.step(40)
    b.step(40)
    assert a.counters()["rebuilds"] == b.counters()["rebuilds"]
    for k in range(system.n_replicas):
        ya, _, ra, ta = a.get_state(k)
        yb, _, rb, tb = b.get_state(k)
        assert ra == rb and ta == tb and np.array_equal(ya, yb)
        ca, sa = a.get_qa_state(k)
        cb, sb = b.get_qa_state(k)
        assert np.array_equal(ca, cb) and np.array_equal(sa, sb)
        for name in _abi.LISTS:
            la, lb = a.get_list(k, name), b.get_list(k, name)
            assert np.array_equal(la[0], lb[0]) and np.array_equal(la[1], lb[1]), name


@pytest.mark.parametrize("lanes", [4, 16])
def test_force_lanes_variants(c2_small, lanes):
    """The 4- and 16-lanes-per-bead variants of the force kernel (the
    autotuner's other choices): per-term forces and energies against the
    oracle on an evolved, perturbed periodic conformation."""
    params, system = c2_small
    params = _abi.XmdParams.from_buffer_copy(params)
    params.force_lanes = lanes
    eng, ora = Engine(params, system), Oracle(params, system)
    ora.step(300)
    y, _, rng, t = ora.get_state(0)
    cont, sync = ora.get_qa_state(0)
    eng.set_state(0, y, rng, t)
    eng.set_qa_state(0, cont, sync)
    ora.set_state(0, y, rng, t)
    r = y[0] + np.random.default_rng(6).normal(scale=0.3 * 0.2, size=y[0].shape)
    eng.set_positions(0, r)
    ora.set_positions(0, r)
    compare_terms(eng, ora)


def test_autotune_keeps_state():
    """The autotuner times whole steps from a copy of the engine's state and
    restores it: the state is unchanged (positions, Gear derivatives,
    generator, time) and the trajectory stays in parity with the oracle
    afterwards."""
    system = synth.config_system(2, n_batch=2, chains=6, horizon_steps=60)
    params = synth.config_params(2)
    eng = Engine(params, system)
    ora = Oracle(params, system)
    eng.step(10)
    ora.step(10)
    before = [eng.get_state(k) for k in range(2)]
    lpb, ms = eng.autotune(reps=3)
    assert lpb in (4, 8, 16) and len(ms) == 3 and all(m > 0 for m in ms)
    # the default tile unless another is more than 3 % faster
    assert lpb == (8 if ms[1] <= 1.03 * min(ms) else (4, 8, 16)[int(np.argmin(ms))])
    for k in range(2):
        y, _, rng, t = eng.get_state(k)
        assert np.array_equal(y, before[k][0]) and rng == before[k][2] and t == before[k][3]
    eng.step(10)
    ora.step(10)
    for k in range(2):
        y, _, rng, t = eng.get_state(k)
        yo, _, rngo, to = ora.get_state(k)
        assert rng == rngo and t == to and np.abs(y - yo).max() < 1e-9


def degenerate_system():
    """Chains of 1..5 beads beside two 30-bead chains, once in a periodic box
    and once in an open box (two replicas of one engine): every boundary
    case of the bonded windows (no prev/next, tether only, angle without
    dihedral) and of PID (needs both neighbours) and QA (n, h = 0 at ends)."""
    rng = np.random.default_rng(21)
    lens = [1, 2, 3, 4, 5, 30, 30]
    chains = [synth.random_chain(m, rng) for m in lens]
    L = 70.0
    placed = synth.place_chains([c[0] for c in chains], L, rng)
    reps = [dict(chains=[(p, c[1]) for p, c in zip(placed, chains)], box_A=(L, L, L), temperature=0.35, seed=0x5EED0101),
            dict(chains=[(p, c[1]) for p, c in zip(placed, chains)], box_A=(0, 0, 0), temperature=0.45, seed=0x5EED0102)]
    return synth.make_system(reps)


@pytest.mark.parametrize("prec", [_abi.PREC_FP64, _abi.PREC_MIXED])
def test_degenerate_chains(prec):
    system = degenerate_system()
    params = synth.config_params(2, precision=prec)
    synth.apply_terms(params, {"snd": True})
    pf = synth.config_params(2)
    synth.apply_terms(pf, {"snd": True})
    eng, ora = Engine(params, system), Oracle(pf, system)
    ora.rebuild_lists()
    for k in range(2):
        lists_equal(eng, ora, k)
    # off the native structure (there the bonded forces are round-off)
    r = synth.perturbed_positions(system, seed=23)
    for k in range(2):
        a, b = system.replica_offset[k], system.replica_offset[k + 1]
        eng.set_positions(k, r[:, a:b])
        ora.set_positions(k, r[:, a:b])
    ora.rebuild_lists()
    for k in range(2):
        lists_equal(eng, ora, k)
    if prec == _abi.PREC_FP64:
        compare_terms(eng, ora, rep_count=2)
    else:
        F, _ = eng.eval_forces()
        Fo, _ = ora.eval_forces()
        assert rel_rms(F, Fo) < 1e-5
    eng.step(30)
    ora.step(30)
    for k in range(2):
        y, _, rng_e, t = eng.get_state(k)
        yo, _, rng_o, to = ora.get_state(k)
        assert rng_e == rng_o and t == to
        assert np.abs(y[0] - yo[0]).max() < (1e-9 if prec == _abi.PREC_FP64 else 1e-4)
        if prec == _abi.PREC_FP64:
            lists_equal(eng, ora, k)


@pytest.mark.parametrize("kind,d", [("ss", 6.0), ("ss", 7.0), ("bb", 4.8)])
def test_qa_timeline_engine(kind, d):
    import qa_scenarios as Q
    eng = Engine(Q.qa_params(), Q.qa_system(kind, d))
    got = Q.run_timeline(eng, kind, d)
    Q.check_timeline(got, Q.expected_timeline(kind, d))


@pytest.mark.parametrize("prec", [_abi.PREC_FP64, _abi.PREC_MIXED])
def test_observables_match_host(prec):
    """Per-replica observables computed on the GPU (two stream groups over
    four boxes) against numpy on the same state: Rg, K, Q within FP64
    reduction order, W (closed-form 3x3 eigenvalues vs SVD) within 1e-9 of
    the squared singular values, V equal to the summed per-bead energies."""
    from xmd_amd.observables import gyration_radius, asphericity, kinetic_energy
    system = synth.config_system(2, n_batch=4, chains=6, chain_len=80)
    params = synth.config_params(2)
    params.precision = prec
    eng = Engine(params, system)
    eng.step(300)
    obs = eng.observables(1.5)
    V = eng.energy()
    for k in range(system.n_replicas):
        y, v, _, t = eng.get_state(k)
        r = y[0]
        sl = system.replica_slice(k)
        m = (system.nat_i1 >= sl.start) & (system.nat_i1 < sl.stop)
        i1, i2 = system.nat_i1[m] - sl.start, system.nat_i2[m] - sl.start
        d = np.linalg.norm(r[:, i1] - r[:, i2], axis=0)
        act = int((d < 1.5 * 2.0 ** (-1.0 / 6.0) * system.nat_dist[m]).sum())
        K = kinetic_energy(v)
        s2 = (np.linalg.svd(r - r.mean(axis=1, keepdims=True), compute_uv=False) ** 2).sum()
        assert obs[k, 0] == t
        assert abs(obs[k, 1] - V[k]) <= 1e-12 * max(1.0, abs(V[k]))
        assert abs(obs[k, 2] - K) <= 1e-12 * K
        assert abs(obs[k, 3] - (V[k] + K)) <= 1e-12 * max(1.0, abs(V[k]) + K)
        assert abs(obs[k, 4] - gyration_radius(r)) <= 1e-12 * gyration_radius(r)
        assert abs(obs[k, 5] - asphericity(r)) <= 1e-9 * s2
        assert obs[k, 7] == m.sum() and abs(obs[k, 6] - act) <= 1, (obs[k, 6], act)
    eng.close()


def ss_system():
    """A 40-bead chain with three disulfide bonds: two within reach, one
    beyond 2 r0 + pad."""
    import dataclasses
    base = synth.config_system(1, chain_len=40)
    r = base.r0.copy()
    r[:, 2] = r[:, 20] + np.array([1.1, 0.3, -0.2])
    r[:, 30] = r[:, 5] + np.array([0.0, 1.4, 0.5])
    r[:, 39] = r[:, 10] + np.array([6.0, 0.0, 0.0])
    system = dataclasses.replace(base, r0=r, ss_i1=np.array([2, 5, 10], dtype=np.int32),
                                 ss_i2=np.array([20, 30, 39], dtype=np.int32))
    params = synth.config_params(1)
    params.nat_ss_enabled = 1
    return params, system


@pytest.mark.parametrize("prec", [_abi.PREC_FP64, _abi.PREC_MIXED])
def test_nat_ss_bonds(prec):
    params, system = ss_system()
    params.precision = prec
    eng = Engine(params, system)
    params.precision = _abi.PREC_FP64
    ora = Oracle(params, system)
    ora.rebuild_lists()
    lists_equal(eng, ora, 0)
    assert set(zip(*eng.get_list(0, "nat_ss"))) == {(2, 20), (5, 30)}
    if prec == _abi.PREC_FP64:
        compare_terms(eng, ora)
        eng.step(20)
        ora.step(20)
        y, _, rng, t = eng.get_state(0)
        yo, _, rngo, to = ora.get_state(0)
        assert rng == rngo and t == to
        assert np.abs(y - yo).max() < 1e-9
    else:
        k = _abi.TERM["nat_ss"]
        F, V = eng.eval_forces(1 << k)
        Fo, Vo = ora.eval_forces(1 << k)
        assert rel_rms(F, Fo) < 1e-12 and abs(V[0, k] - Vo[0, k]) < 1e-12 * abs(Vo[0, k])
    eng.close()


def test_checkpoint_restart_engine(tmp_path):
    """Engine checkpoints: restart bit-identical to the saved run after the
    same rebuild; the same checkpoint loaded into the oracle continues at
    the FP64 trajectory bar."""
    from xmd_amd import checkpoint
    system = synth.config_system(2, n_batch=2, chains=4, chain_len=60)
    params = synth.config_params(2)
    a = Engine(params, system)
    a.step(60)
    path = str(tmp_path / "ckpt.npz")
    checkpoint.save(a, path)
    checkpoint.load(a, path)
    b = Engine(params, system)
    checkpoint.load(b, path)
    ora = Oracle(params, system)
    checkpoint.load(ora, path)
    a.step(20)
    b.step(20)
    ora.step(20)
    for r in range(2):
        ya, _, ra, ta = a.get_state(r)
        yb, _, rb, tb = b.get_state(r)
        yo, _, ro, to = ora.get_state(r)
        assert ra == rb == ro and ta == tb == to
        assert np.array_equal(ya, yb)
        assert np.abs(ya - yo).max() < 1e-8
        assert np.array_equal(a.get_qa_state(r)[1], b.get_qa_state(r)[1])
    a.close()
    b.close()


@pytest.mark.parametrize("prec", [_abi.PREC_FP64, _abi.PREC_MIXED])
def test_walls_and_afm(prec, tmp_path):
    """Solid and LJ attractive walls and both AFM tips against the oracle:
    per-term forces (FP64 in both forms), 40-step trajectories with the
    wall pairs' states bit-equal, and a checkpoint carrying them."""
    from test_oracle_pins import walls_afm_system
    from xmd_amd import checkpoint
    system = walls_afm_system(synth.config_system(1, chain_len=40))
    params = synth.config_params(1)
    params.solid_walls_enabled = params.lj_walls_enabled = 1
    params.vel_afm_enabled = params.force_afm_enabled = 1
    params.lj_walls_phase_duration = 0.05
    ora = Oracle(params, system)
    params.precision = prec
    eng = Engine(params, system)
    for name in ("solid_wall", "lj_wall", "vel_afm", "force_afm"):
        k = _abi.TERM[name]
        F, V = eng.eval_forces(1 << k)
        Fo, Vo = ora.eval_forces(1 << k)
        assert rel_rms(F, Fo) < 1e-10, name
        assert abs(V[0, k] - Vo[0, k]) <= 1e-10 * max(1.0, abs(Vo[0, k])), name
    if prec == _abi.PREC_MIXED:
        eng.close()
        return
    # pairs join at once, saturate after 10 steps, break and free later:
    # at step 20 some are formed and some breaking, at step 40 all free
    states = []
    for _ in range(2):
        eng.step(20)
        ora.step(20)
        y, _, rng, t = eng.get_state(0)
        yo, _, rngo, to = ora.get_state(0)
        assert rng == rngo and t == to
        assert np.abs(y - yo).max() < 1e-9
        st, jo, rt = eng.get_wall_state(0)
        sto, joo, rto = ora.get_wall_state(0)
        assert np.array_equal(st, sto) and np.array_equal(rt, rto) and np.abs(jo - joo).max() < 1e-9
        states.append(st.copy())
        if len(states) == 1:
            ckpt = str(tmp_path / "ckpt.npz")
            checkpoint.save(eng, ckpt)
    assert {1, 2} <= set(states[0].tolist())
    eng2 = Engine(params, system)
    checkpoint.load(eng2, ckpt)
    assert np.array_equal(eng2.get_wall_state(0)[0], states[0])
    eng.close()
    eng2.close()


@pytest.mark.parametrize("kind", ["ortho", "slab"])
@pytest.mark.parametrize("prec", [_abi.PREC_FP64, _abi.PREC_MIXED])
def test_odd_boxes(kind, prec):
    """A non-cubic periodic box (cell lists of unequal sides) and a slab
    periodic in x and y only (the engine's all-pairs path with a partial
    minimum image): lists bit-exact in both precisions, forces and a
    10-step trajectory with QA state at the usual bars."""
    from test_oracle_pins import odd_box_system
    system = odd_box_system(kind)
    print("box", kind, system.box[0], flush=True)
    eng = Engine(synth.config_params(2, precision=prec), system)
    ora = Oracle(synth.config_params(2), system)
    r = synth.perturbed_positions(system, 5)
    eng.set_positions(0, r)
    ora.set_positions(0, r)
    eng.rebuild_lists()
    ora.rebuild_lists()
    lists_equal(eng, ora, 0)
    if prec == _abi.PREC_FP64:
        compare_terms(eng, ora)
    else:
        F, _ = eng.eval_forces()
        Fo, _ = ora.eval_forces()
        assert rel_rms(F, Fo) < 1e-5
    eng.step(10)
    ora.step(10)
    y, _, rng, t = eng.get_state(0)
    yo, _, rngo, to = ora.get_state(0)
    assert rng == rngo and t == to
    assert np.abs(y[0] - yo[0]).max() < (1e-9 if prec == _abi.PREC_FP64 else 1e-4)
    if prec == _abi.PREC_FP64:
        ce, se = eng.get_qa_state(0)
        co, so = ora.get_qa_state(0)
        assert len(ce) == len(co) and np.array_equal(se, so)
        lists_equal(eng, ora, 0)
    eng.close()


def _variant_params(variant, precision=_abi.PREC_FP64):
    p = synth.config_params(2, precision=precision)
    if variant in ("rel_es", "both_es"):
        p.rel_es_enabled = 1
        p.const_es_enabled = int(variant == "both_es")
    elif variant == "algebraic":
        p.pid_lambda_version = _abi.LAMBDA_ALGEBRAIC
    elif variant == "no_sink":
        p.ss_sinking = 0
    return p


@pytest.mark.parametrize("variant", ["rel_es", "both_es", "algebraic", "no_sink"])
def test_term_variants(variant, c2_small):
    """Parameter variants of the periodic full-term box: relative
    Debye-Hueckel alone and beside the constant one (D31), the algebraic
    PID lambda, plain LJ for the side-side variants (no sink): lists,
    per-term forces and 10 steps against the oracle in FP64, forces within
    1e-5 in mixed precision."""
    _, system = c2_small
    print("variant", variant, flush=True)
    eng, ora = Engine(_variant_params(variant), system), Oracle(_variant_params(variant), system)
    emx = Engine(_variant_params(variant, _abi.PREC_MIXED), system)
    r = synth.perturbed_positions(system, 9)
    for e in (eng, ora, emx):
        e.set_positions(0, r)
        e.rebuild_lists()
    lists_equal(eng, ora, 0)
    compare_terms(eng, ora)
    F, _ = emx.eval_forces()
    Fo, _ = ora.eval_forces()
    assert rel_rms(F, Fo) < 1e-5
    eng.step(10)
    ora.step(10)
    y, _, rng, t = eng.get_state(0)
    yo, _, rngo, to = ora.get_state(0)
    assert rng == rngo and t == to and np.abs(y - yo).max() < 1e-9
    ce, se = eng.get_qa_state(0)
    co, so = ora.get_qa_state(0)
    assert len(ce) == len(co) and np.array_equal(se, so)
    eng.close()
    emx.close()


def test_c0_cutoff_override():
    """C0: the 60-bead chain with the list radius fixed at 18 A + 2 A pad
    (max_cutoff_override) instead of the term cutoffs: lists, forces and a
    200-step trajectory against the oracle."""
    system = synth.config_system(0)
    params = synth.config_params(0)
    assert params.max_cutoff_override > 0
    eng, ora = Engine(params, system), Oracle(params, system)
    r = synth.perturbed_positions(system, 3)
    eng.set_positions(0, r)
    ora.set_positions(0, r)
    eng.rebuild_lists()
    ora.rebuild_lists()
    lists_equal(eng, ora, 0)
    compare_terms(eng, ora)
    # rebuilds during the steps (the engine also counts the one made when
    # it was created)
    rb0, rbo0 = eng.counters()["rebuilds"], ora.counters()["rebuilds"]
    eng.step(200)
    ora.step(200)
    y, _, rng, t = eng.get_state(0)
    yo, _, rngo, to = ora.get_state(0)
    assert rng == rngo and t == to and np.abs(y - yo).max() < 1e-8
    assert eng.counters()["rebuilds"] - rb0 == ora.counters()["rebuilds"] - rbo0 > 0
    eng.close()


def test_verlet_list_complete_between_rebuilds():
    """Between rebuilds the Verlet list must hold every pair closer than
    the list cutoff (nl::verify's skin criterion, verify.cpp:10-21): a
    4-box batch of the periodic full-term workload (10k beads each) checked
    against a k-d tree every 100 of 2000 steps."""
    from scipy.spatial import cKDTree
    batch = synth.config_system(2, n_batch=4, horizon_steps=2000)
    p = synth.config_params(2, precision=_abi.PREC_MIXED)
    ss = np.array(p.ss_rmax)
    mc = max(p.pauli_r_excl, (2.0 * batch.nat_dist).max(), 2.0 * p.const_es_screen,
             max(p.lj_bb_rmin, p.lj_bs_rmin, ss.max()), max(2 * p.lj_bb_rmin, 2 * p.lj_bs_rmin, (2 * ss).max()))
    eng = Engine(p, batch)
    rebuilds = eng.counters()["rebuilds"]
    for step in range(100, 2001, 100):
        eng.step(100)
        for k in range(batch.n_replicas):
            sl = batch.replica_slice(k)
            y = eng.get_state(k)[0][0].T
            L = batch.box[k]
            tree = cKDTree(y - L * np.floor(y / L), boxsize=L * (1 + 1e-12))
            pr = tree.query_pairs(mc * (1 - 1e-9), output_type="ndarray")
            i, j = pr.min(axis=1), pr.max(axis=1)
            chain = np.searchsorted(batch.chain_offset, np.arange(sl.start, sl.stop), side="right") - 1
            keep = ~((chain[i] == chain[j]) & (j - i < 3))
            a, b = eng.get_list(k, "nl_all")
            have = set(zip(a.tolist(), b.tolist()))
            missing = [q for q in zip(i[keep].tolist(), j[keep].tolist()) if q not in have]
            assert not missing, (step, k, missing[:5])
    assert eng.counters()["rebuilds"] > rebuilds
    eng.close()
