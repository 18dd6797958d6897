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


def gpu_sample(eng, nrep, nsteps, every, burn):
    """Per-sample, per-replica (V, K, Rg, native fraction, QA contacts),
    the first four from the engine's GPU observables."""
    eng.step(burn)
    out = []
    for _ in range(nsteps // every):
        eng.step(every)
        obs = eng.observables(1.5)
        qa = [len(eng.get_qa_state(k)[0]) for k in range(nrep)]
        out.append(np.column_stack([obs[:, 1], obs[:, 2], obs[:, 4], obs[:, 6] / np.maximum(obs[:, 7], 1), qa]))
    return np.array(out)  # (samples, replicas, 5)


def test_mixed_precision_samples_the_fp64_ensemble():
    """The mixed-precision engine against the FP64 engine on the full term
    set (Go, Pauli, DH, PID, QA, bonded): 32 boxes of 4 x 60 beads at
    T = 0.30 .. 0.50, same initial states and noise streams; trajectories
    decorrelate, so per-replica time averages are compared as paired
    differences within 4 standard errors."""
    from xmd_amd import _abi
    nrep = 32
    system = synth.config_system(2, n_batch=nrep, chains=4, chain_len=60, horizon_steps=6000)
    params = synth.config_params(2)
    params.precision = _abi.PREC_FP64
    e64 = Engine(params, system)
    params.precision = _abi.PREC_MIXED
    emx = Engine(params, system)
    a = gpu_sample(emx, nrep, 4000, 100, 1000)
    b = gpu_sample(e64, nrep, 4000, 100, 1000)
    assert b[..., 4].mean() > 1  # QA contacts form
    for q, name in enumerate(["V", "K", "Rg", "Q", "QA contacts"]):
        d = a[:, :, q].mean(axis=0) - b[:, :, q].mean(axis=0)
        se = d.std(ddof=1) / np.sqrt(nrep) + 1e-12 * abs(b[:, :, q].mean())
        assert abs(d.mean()) < 4 * se, (name, a[:, :, q].mean(), b[:, :, q].mean(), se)
    emx.close()
    e64.close()


def host_sample(hosts, systems, nsteps, every, burn):
    """Per-sample, per-replica (Rg, K, native fraction, QA contacts) of a
    list of drivers (engine or oracles, one system each), stepped side by
    side on threads (the oracle's calls release the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=len(hosts)) as ex:
        list(ex.map(lambda h: h.step(burn), hosts))
        out = []
        for _ in range(nsteps // every):
            list(ex.map(lambda h: h.step(every), hosts))
            row = []
            for h, sy in zip(hosts, systems):
                for k in range(sy.n_replicas):
                    row.append(list(replica_observables(h, sy, k)) + [len(h.get_qa_state(k)[0])])
            out.append(row)
    return np.array(out)  # (samples, replicas, 4)


def test_full_term_set_ensemble_matches_oracle():
    """The FP64 engine against the oracle on the full term set (Go, Pauli,
    Debye-Hueckel, PID, QA, tether, chirality, native angles and
    dihedrals): 32 boxes of 4 x 60 beads at C2's density and T = 0.30 ..
    0.50, the same initial states and noise streams; trajectories
    decorrelate, so per-replica time averages of Rg, K, native fraction and
    QA contacts are compared as paired differences within 4 standard
    errors (reference: stats/gyration.cpp:4-18, report_structure.cpp:81-85,
    dynamics/lang_pc.cpp:5-41)."""
    nrep, ngroups = 32, 8
    system = synth.config_system(2, n_batch=nrep, chains=4, chain_len=60, horizon_steps=6000)
    params = synth.config_params(2)
    eng = Engine(params, system)
    groups = [list(range(g * nrep // ngroups, (g + 1) * nrep // ngroups)) for g in range(ngroups)]
    subs = [synth.subset(system, ks) for ks in groups]
    oras = [Oracle(params, sy) for sy in subs]
    a = host_sample([eng], [system], 4000, 200, 1000)
    b = host_sample(oras, subs, 4000, 200, 1000)
    assert b[..., 3].mean() > 1  # QA contacts form
    for q, name in enumerate(["Rg", "K", "Q", "QA contacts"]):
        d = a[:, :, q].mean(axis=0) - b[:, :, q].mean(axis=0)
        se = d.std(ddof=1) / np.sqrt(nrep) + 1e-12 * abs(b[:, :, q].mean())
        print(name, a[:, :, q].mean(), b[:, :, q].mean(), d.mean(), se, flush=True)
        assert abs(d.mean()) < 4 * se, (name, a[:, :, q].mean(), b[:, :, q].mean(), se)
    eng.close()
