"""Neighbour rows that outgrow their width (the reference's lists are
vectors that grow, xmd/src/nl/legacy_update.cpp:28-59): the engine widens
the rows, rebuilds the replica that needed it from the state it halted at
and steps it alone back to the others.  Every kernel exits normally: the
overflow is a flag the row fill sets and the host reads after the step."""
import numpy as np
import pytest

from xmd_amd import synth
from xmd_amd.engine import Engine
from test_gpu_parity import contacts_equal

pytestmark = pytest.mark.gpu


def test_rows_grow_like_a_wider_engine():
    """Rows exactly as wide as the longest first row (row_width 1) against
    rows of 1024 entries: after 3000 steps every replica's state, QA state
    and the rebuild count are equal bit for bit, and the rows have grown."""
    system = synth.config_system(2, n_batch=3, chains=8, chain_len=60, horizon_steps=4000)
    tight, wide = synth.config_params(2), synth.config_params(2)
    tight.row_width, wide.row_width = 1, 1024
    a, b = Engine(tight, system), Engine(wide, system)
    cap0 = a.counters()["nl_capacity"]
    for _ in range(6):
        a.step(500)
        b.step(500)
    ca, cb = a.counters(), b.counters()
    print("capacity", cap0, "->", ca["nl_capacity"], "of", cb["nl_capacity"], flush=True)
    assert ca["nl_capacity"] > cap0
    assert ca["steps"] == cb["steps"] == 3000 and ca["rebuilds"] == cb["rebuilds"]
    for k in range(system.n_replicas):
        ya, _, ra, ta = a.get_state(k)
        yb, _, rb, tb = b.get_state(k)
        assert ra == rb and ta == tb, k
        assert np.array_equal(ya, yb), k
        contacts_equal(a, b, rep=k)


def test_free_pairs_grow_like_a_larger_engine():
    """QA free-pair regions of 256 entries per replica against the automatic
    capacity: the first rebuild already outgrows them, and the chains'
    collapse adds more than the quarter of headroom each growth leaves, so
    they grow again while stepping (the count of free pairs passes what the
    first growth could hold); after 3000 steps every replica's state, QA
    contacts, free pairs and rebuild count equal the larger engine's."""
    system = synth.config_system(2, n_batch=3, chains=8, chain_len=60, horizon_steps=4000)
    tight, auto = synth.config_params(2), synth.config_params(2)
    tight.fp_capacity = 256
    a, b = Engine(tight, system), Engine(auto, system)
    nrep = system.n_replicas
    free = lambda e, k: e.get_list(k, "qa_free")
    n0 = [len(free(a, k)[0]) for k in range(nrep)]
    first = [max(256 * -(-(n + n // 4) // 256), 512) for n in n0]  # capacity after the first growth
    peak = list(n0)
    for _ in range(6):
        a.step(500)
        b.step(500)
        peak = [max(p, len(free(a, k)[0])) for k, p in enumerate(peak)]
    print("free pairs at creation", n0, "first growth to", first, "peak", peak, flush=True)
    assert min(n0) > 256
    assert any(p > c for p, c in zip(peak, first))
    ca, cb = a.counters(), b.counters()
    assert ca["steps"] == cb["steps"] == 3000 and ca["rebuilds"] == cb["rebuilds"]
    for k in range(nrep):
        ya, _, ra, ta = a.get_state(k)
        yb, _, rb, tb = b.get_state(k)
        assert ra == rb and ta == tb, k
        assert np.array_equal(ya, yb), k
        contacts_equal(a, b, rep=k)
        for x, y in zip(free(a, k), free(b, k)):
            assert np.array_equal(x, y), k
    a.close()
    b.close()


def test_contact_compaction_keeps_results():
    """QA contact sets of 512 entries, compacted in order whenever removed
    contacts take their extent past 256 (the boxes hold up to ~240 live
    contacts per replica by the end), against the automatic capacity that
    never compacts here: after 3000 steps every replica's state, contact set
    and rebuild count are equal bit for bit."""
    system = synth.config_system(2, n_batch=3, chains=8, chain_len=60, horizon_steps=4000)
    tight, auto = synth.config_params(2), synth.config_params(2)
    tight.ct_capacity = 512
    a, b = Engine(tight, system), Engine(auto, system)
    for _ in range(6):
        a.step(500)
        b.step(500)
    ca, cb = a.counters(), b.counters()
    assert ca["steps"] == cb["steps"] == 3000 and ca["rebuilds"] == cb["rebuilds"]
    n = 0
    for k in range(system.n_replicas):
        ya, _, ra, ta = a.get_state(k)
        yb, _, rb, tb = b.get_state(k)
        assert ra == rb and ta == tb, k
        assert np.array_equal(ya, yb), k
        n += contacts_equal(a, b, rep=k)
    print("contacts at the end", n, flush=True)
    assert n > 3 * 128  # above half the capacity on average: compaction ran
    a.close()
    b.close()
