"""The plain-C client (tests/c_client/xmd_c_client.c) on the GPU: its
engine, filled and driven from C, gives bit for bit the forces and the
200-step trajectory that the ctypes path gives for the same structs, and
the oracle's trajectory within the FP64 bars."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from xmd_amd import _abi
from xmd_amd.engine import Engine
from xmd_amd.synth import System
from oracle.oracle import Oracle
from test_c_client import client

pytestmark = pytest.mark.gpu


def read_run(path):
    b = open(path, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(b, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a.copy()

    n, nch, nn, nterms = take(np.int32, 4)
    params = _abi.XmdParams.from_buffer_copy(b[off:off + C.sizeof(_abi.XmdParams)])
    off += C.sizeof(_abi.XmdParams)
    seed = take(np.uint64, 1)
    temp = take(np.float64, 1)
    r0 = take(np.float64, 3 * n).reshape(3, n)
    atype = take(np.int8, n)
    i1, i2, dist = take(np.int32, nn), take(np.int32, nn), take(np.float64, nn)
    chir, chir_f = take(np.float64, n), take(np.float64, n)
    F, Vt = take(np.float64, 3 * n).reshape(3, n), take(np.float64, nterms)
    y = take(np.float64, 18 * n).reshape(6, 3, n)
    rng, t, V = take(np.uint64, 1)[0], take(np.float64, 1)[0], take(np.float64, 1)[0]
    counters = take(np.int64, 5)
    assert off == len(b)
    nan = np.full(n, np.nan)
    system = System(replica_offset=np.array([0, n], np.int32),
                    chain_offset=np.arange(nch + 1, dtype=np.int32) * (n // nch), atype=atype, r0=r0,
                    box=np.zeros((1, 3)), temperature=temp, rng_state=seed, nat_i1=i1, nat_i2=i2, nat_dist=dist,
                    tether_len=nan, nat_theta=nan, nat_phi=nan, nat_chir=chir, nat_chir_factor=chir_f)
    return params, system, dict(F=F, Vt=Vt, y=y, rng=int(rng), t=t, V=V, counters=counters)


def test_c_client_matches_ctypes_and_oracle(tmp_path):
    out = str(tmp_path / "run.bin")
    print("c client run", flush=True)
    res = subprocess.run([client(), "run", out], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    params, system, c = read_run(out)
    eng = Engine(params, system)
    F, V = eng.eval_forces(0xFFFFFFFF)  # the mask the C client passes
    assert np.array_equal(F, c["F"]) and np.array_equal(V[0], c["Vt"])
    eng.step(200)
    y, _, rng, t = eng.get_state(0)
    assert rng == c["rng"] and t == c["t"] and np.array_equal(y, c["y"])
    assert eng.energy()[0] == c["V"]
    eng.close()
    ora = Oracle(params, system)
    Fo, _ = ora.eval_forces()
    assert np.sqrt(((c["F"] - Fo) ** 2).sum() / (Fo ** 2).sum()) < 1e-10
    ora.step(200)
    yo, _, rngo, to = ora.get_state(0)
    assert rngo == c["rng"] and to == c["t"]
    assert np.abs(yo - c["y"]).max() < 1e-8
