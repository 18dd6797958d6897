"""Long-run ensemble averages of the engine against the oracle (C0 force
field: Go, Pauli, tether, chirality; 32 replicas of 60 beads).  Both start
from identical states with the same noise stream; trajectories decorrelate,
so only averages are compared, within 4 standard errors."""
import numpy as np
import pytest

from xmd_amd import synth
from xmd_amd.engine import Engine
from xmd_amd.observables import replica_observables
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def sample(host, system, nrep, nsteps, every, burn):
    host.step(burn)
    out = []
    for _ in range(nsteps // every):
        host.step(every)
        out.append([replica_observables(host, system, k) for k in range(nrep)])
    return np.array(out)  # (samples, replicas, 3)


def test_ensemble_averages_match_oracle():
    nrep = 32
    system = synth.config_system(0, n_batch=nrep, horizon_steps=6000)
    params = synth.config_params(0)
    eng, ora = Engine(params, system), Oracle(params, system)
    a = sample(eng, system, nrep, 4000, 200, 1000)
    b = sample(ora, system, nrep, 4000, 200, 1000)
    for q, name in enumerate(["Rg", "K", "Q"]):
        ra, rb = a[:, :, q].mean(axis=0), b[:, :, q].mean(axis=0)  # per-replica time averages
        diff = ra.mean() - rb.mean()
        se = np.sqrt((ra.var(ddof=1) + rb.var(ddof=1)) / nrep) + 1e-12 * abs(rb.mean())
        assert abs(diff) < 4 * se + 1e-9, (name, ra.mean(), rb.mean(), se)
