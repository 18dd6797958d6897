"""Error paths of the engine, forced from the host; every kernel still
exits normally (overflows are flagged and truncated, non-finite values are
detected after the step)."""
import numpy as np
import pytest

from xmd_amd import synth, _abi
from xmd_amd._host import XmdError
from xmd_amd.engine import Engine, first_zero_log_draw

pytestmark = pytest.mark.gpu


def test_row_overflow_is_a_capacity_error():
    # 1200 beads squeezed into a 12 A cube: every row holds ~1199 partners,
    # more than the widest row an engine allows (1024)
    system = synth.config_system(2, chains=40, chain_len=30)
    eng = Engine(synth.config_params(2, precision=_abi.PREC_MIXED), system)
    eng.step(2)
    rng = np.random.default_rng(3)
    r = system.r0[:, :1] + rng.uniform(0, 12 * synth.ANGSTROM, size=system.r0.shape)
    eng.set_positions(0, r)
    with pytest.raises(XmdError) as e:
        eng.step(1)
    assert e.value.status == _abi.ERR_CAPACITY and "neighbour-list capacity" in str(e.value)


def test_zero_uniform_draw_reports_infinite_noise():
    # seed 0x5EED0007 feeds a zero uniform draw to log() in the Langevin
    # noise of a 10k-bead box (rand_gen.h:45-46: the reference's noise is
    # infinite there); the engine reports it after that step
    seed = 0x5EED0007
    system = synth.config_system(2, seed=seed)
    hit = first_zero_log_draw(seed, system.n_beads, 1000)
    assert hit is not None
    step, bead = hit
    eng = Engine(synth.config_params(2, precision=_abi.PREC_MIXED), system)
    eng.step(step)
    y, _, _, _ = eng.get_state(0)
    assert np.isfinite(y).all()
    with pytest.raises(XmdError) as e:
        eng.step(1)
    assert e.value.status == _abi.ERR_NONFINITE and "zero uniform draw" in str(e.value)


# rand_gen state whose next output is 1 (multiplier inverse mod 2^64): the
# first draw of the next step has zero upper bits, a zero uniform for log()
_M = 0xD989BCACC137DCD5
ZERO_NEXT = pow(_M, -1, 1 << 64)


def test_halted_replica_leaves_the_others_unchanged():
    system = synth.config_system(3, n_replicas=3)
    p = synth.config_params(3, precision=_abi.PREC_FP64)
    a, b = Engine(p, system), Engine(p, system)
    a.step(10)
    b.step(10)
    y1, _, _, t1 = a.get_state(1)
    a.set_state(1, y1, ZERO_NEXT, t1)
    yb, _, rb, tb = b.get_state(1)
    b.set_state(1, yb, rb, tb)  # the same forced list rebuild as in a
    with pytest.raises(XmdError) as e:
        a.step(5)
    assert e.value.status == _abi.ERR_NONFINITE and "replica 1:" in str(e.value)
    assert a.replica_status(1)[0] == _abi.ERR_NONFINITE and a.replica_status(0)[0] == 0
    assert a.get_state(1)[3] == t1  # halted: time does not advance
    a.step(3)  # reported once; the halted replica stays put
    assert a.get_state(1)[3] == t1
    with pytest.raises(XmdError) as e:
        a.eval_forces()
    assert e.value.status == _abi.ERR_STATE
    b.step(8)
    # reset replica 1 of a to b's state; b takes the same forced rebuild
    ys, _, rs, ts = b.get_state(1)
    a.set_state(1, ys, rs, ts)
    b.set_state(1, ys, rs, ts)
    assert a.replica_status(1)[0] == 0
    a.step(7)
    b.step(7)
    for k in range(3):
        ya, _, ra, ta = a.get_state(k)
        yb, _, rb, tb = b.get_state(k)
        assert ra == rb and ta == tb, k
        assert np.array_equal(ya, yb), k


def test_autotune_with_a_halted_replica_changes_nothing():
    """The autotuner steps every replica from a copy of the state and
    restores it: a halted replica stays halted (its error still reported by
    xmd_replica_status), and every replica's state equals that of a twin
    engine that did not tune."""
    system = synth.config_system(3, n_replicas=3)
    p = synth.config_params(3, precision=_abi.PREC_MIXED)
    a, b = Engine(p, system), Engine(p, system)
    for e in (a, b):
        e.step(10)
        y1, _, _, t1 = e.get_state(1)
        e.set_state(1, y1, ZERO_NEXT, t1)
        with pytest.raises(XmdError):
            e.step(2)
    lanes, ms = a.autotune(reps=2)
    assert lanes in (4, 8, 16) and all(m > 0 for m in ms)
    assert a.replica_status(1)[0] == _abi.ERR_NONFINITE and a.replica_status(0)[0] == 0
    for k in range(3):
        ya, _, ra, ta = a.get_state(k)
        yb, _, rb, tb = b.get_state(k)
        assert ra == rb and ta == tb
        assert np.array_equal(ya, yb, equal_nan=True)
    if lanes == 8:  # the twins' tiles agree: they step alike from here
        a.step(5)
        b.step(5)
        for k in (0, 2):
            assert np.array_equal(a.get_state(k)[0], b.get_state(k)[0])
