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
