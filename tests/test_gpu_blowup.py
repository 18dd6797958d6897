"""A replica that turns non-finite in the engine does so under the
reference's own dynamics.

Native chains free of the bond-angle cap (synth.MAX_BOND_ANGLE, max_angle
None) give the C2 box near-straight native angles, where the native-dihedral
and PID gradients diverge, and replicas heat up and blow up within a few
thousand steps.  The oracle is handed the engine's full state -- Gear
derivatives, generator, time, QA contacts and budgets -- at a step where the
engine rebuilt its lists, about 200 steps before the engine's replica turns
non-finite: the two then differ only by rounding amplified by the chaotic
dynamics (about tenfold per 25 steps), and the oracle turns non-finite at the
same step (reference: xmd/src/freeform/main.cpp:1028-1076,
dynamics/lang_pc.cpp:5-41).
"""
import numpy as np
import pytest

from xmd_amd import synth
from xmd_amd.engine import Engine
from xmd_amd._host import XmdError
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

NONFINITE = 4  # XMD_ERR_NONFINITE


def _step(e, n):
    try:
        e.step(n)
    except XmdError:
        pass


def _halted(e, k=0):
    return e.replica_status(k)[0] != 0


def test_oracle_follows_engine_into_blowup():
    params = synth.config_params(2)
    batch = synth.config_system(2, n_batch=64, seed=0x5EED2000, horizon_steps=10000, max_angle=None)
    eng = Engine(params, batch)
    first = None
    for _ in range(6):
        _step(eng, 500)
        hit = [k for k in range(batch.n_replicas) if eng.replica_status(k)[0] == NONFINITE]
        if hit:
            first = hit[0]
            break
    eng.close()
    if first is None:
        pytest.skip("no replica of the unrestricted batch turned non-finite within 3000 steps")
    one = synth.config_system(2, n_batch=1, seed=int(batch.rng_state[first]), max_angle=None)
    one.temperature[:] = batch.temperature[first]
    print("replica", first, "seed", int(batch.rng_state[first]), flush=True)

    # the step B whose positions are non-finite, replica alone
    e = Engine(params, one)
    done = 0
    while not _halted(e) and done < 3500:
        _step(e, 50)
        done += 50
    assert _halted(e)
    e.close()
    e = Engine(params, one)
    _step(e, done - 50)
    B = done - 50
    while not _halted(e):
        _step(e, 1)
        B += 1
    e.close()
    # the last list rebuild at least 150 steps before B
    e = Engine(params, one)
    lo = max(0, B - 400)
    _step(e, lo)
    S0 = None
    for s in range(lo + 1, B - 150 + 1):
        r0 = e.counters()["rebuilds"]
        _step(e, 1)
        if e.counters()["rebuilds"] > r0:
            S0 = s
    assert S0 is not None
    e.close()
    e = Engine(params, one)
    _step(e, S0)
    y, _, rng, t = e.get_state(0)
    ct, sync = e.get_qa_state(0)
    ora = Oracle(params, one)
    ora.set_state(0, y, rng, t)
    ora.set_qa_state(0, ct, sync)
    print("hand-over at step", S0, "engine non-finite at step", B, flush=True)
    ora_bad = None
    for s in range(S0 + 1, B + 3):
        if s <= B:
            _step(e, 1)
        ora.step(1)
        ro = ora.positions(0)
        if not np.isfinite(ro).all():
            ora_bad = s
            break
        if s <= B - 10:
            gap = np.abs(ro - e.positions(0)).max()
            if s == S0 + 10:
                assert gap < 1e-10, gap  # the hand-over is exact up to rounding
            # ... and stays within the chaotic drift (a tenth of an angstrom,
            # against 3.8 A bonds: the same trajectory, not another one)
            assert gap < 0.1, (s, gap)
    e.close()
    ora.close()
    print("oracle non-finite at step", ora_bad, flush=True)
    assert ora_bad is not None and abs(ora_bad - B) <= 2, (ora_bad, B)
