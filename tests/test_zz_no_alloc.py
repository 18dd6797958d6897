"""The entry point that allocates no device memory (xmd_observables_noalloc)
against the allocating one (xmd_observables), through the C ABI: two
calls in a row, the second on a smaller engine, and a refusal above the
replica limit."""
import numpy as np
import pytest

from xmd_amd import synth, _abi
from xmd_amd._host import XmdError
from xmd_amd.engine import Engine

pytestmark = pytest.mark.gpu


def test_observables_noalloc_equals_observables():
    for nrep in (40, 7):  # the second call's engine is smaller
        system = synth.config_system(3, n_replicas=nrep)
        eng = Engine(synth.config_params(3, precision=_abi.PREC_MIXED), system)
        eng.step(30)
        a = eng.observables()
        b = eng.observables_noalloc()
        assert a.shape == (nrep, 8) and np.isfinite(a).all()
        assert np.array_equal(a, b), nrep
        eng.close()


def test_observables_noalloc_refuses_more_replicas_than_its_array():
    rng = np.random.default_rng(5)
    n = 4097
    reps = [dict(chains=[synth.random_chain(4, rng)], box_A=(0, 0, 0), temperature=0.35, seed=k + 1)
            for k in range(n)]
    system = synth.make_system(reps)
    eng = Engine(synth.config_params(3, precision=_abi.PREC_MIXED), system)
    with pytest.raises(XmdError) as e:
        eng.observables_noalloc()
    assert e.value.status == _abi.ERR_ARG and "at most 4096" in str(e.value)
    assert eng.observables().shape == (n, 8)
