"""Checkpoints: a run restarted from a checkpoint continues bit for bit
like the saved run after the same forced list rebuild (the reference's
restart semantics); QA contacts, budgets, generator and time included."""
import numpy as np
import pytest

from xmd_amd import synth, checkpoint
from oracle.oracle import Oracle


def _same(a, b):
    for r in range(a.nrep):
        ya, _, ra, ta = a.get_state(r)
        yb, _, rb, tb = b.get_state(r)
        assert ra == rb and ta == tb
        assert np.array_equal(ya, yb)
        ca, sa = a.get_qa_state(r)
        cb, sb = b.get_qa_state(r)
        assert np.array_equal(ca, cb) and np.array_equal(sa, sb)


def test_checkpoint_restart_oracle(tmp_path):
    system = synth.config_system(2, n_batch=2, chains=4, chain_len=60)
    params = synth.config_params(2)
    a = Oracle(params, system)
    a.step(60)
    path = str(tmp_path / "ckpt.npz")
    checkpoint.save(a, path)
    checkpoint.load(a, path)  # the saved run, with the restart's list rebuild
    b = Oracle(params, system)
    checkpoint.load(b, path)
    _same(a, b)
    a.step(40)
    b.step(40)
    _same(a, b)
    assert sum(len(a.get_qa_state(r)[0]) for r in range(a.nrep)) > 0


def test_checkpoint_rejects_other_system(tmp_path):
    system = synth.config_system(0)
    params = synth.config_params(0)
    a = Oracle(params, system)
    path = str(tmp_path / "ckpt.npz")
    checkpoint.save(a, path)
    other = Oracle(params, synth.config_system(0, n_batch=2))
    with pytest.raises(ValueError):
        checkpoint.load(other, path)
