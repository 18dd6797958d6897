"""A run of an input file through the driver (xmd_amd/run.py) on the GPU:
the glut model for 0.2 tau with stats every 0.05 tau.  The reports' first
row holds the oracle's energy and the numpy radius of gyration of the
initial state, and the driver's force evaluations at reporting points
leave the trajectory bit for bit that of the same number of plain steps."""
import os

import numpy as np
import pytest

from xmd_amd import model as M, params as P, run
from xmd_amd.engine import Engine
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

MODELS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models")


def test_driver_run(tmp_path):
    inp = {"model": {"seqfile": "glut/glut.yml"}, "general": {"total time": "0.2 tau"},
           "report stats": {"enabled": True, "path": "out.csv", "exec period": "0.05 tau"},
           "report structure": {"enabled": False}, "export pdb": {"enabled": True, "path": "t.pdb",
                                                                  "exec period": "0.1 tau"}}
    print("driver run", flush=True)
    eng = run.simulate(inp, base_dir=MODELS, out_dir=str(tmp_path))
    y, _, rng, t = eng.get_state(0)
    steps = eng.counters()["steps"]
    eng.close()
    rows = open(tmp_path / "out.csv").read().splitlines()[1:]
    assert len(rows) == 4  # t = 0, 0.05, 0.1, 0.15 (the run stops at t >= 0.2)
    params, system, _ = M.from_input(inp, base_dir=MODELS)
    ora = Oracle(params, system)
    _, V0 = ora.eval_forces()
    first = [float(x) for x in rows[0].split(",")]
    assert first[0] == 0.0 and first[1] == pytest.approx(V0.sum(), rel=1e-6, abs=1e-6)
    r0 = system.r0
    rg = np.sqrt(((r0 - r0.mean(axis=1, keepdims=True)) ** 2).sum(axis=0).mean()) / P.ANGSTROM
    assert first[4] == pytest.approx(rg, rel=1e-6)
    assert open(tmp_path / "t.pdb").read().count("ENDMDL") == 2
    plain = Engine(params, system)
    plain.step(steps)
    yp, _, rngp, tp = plain.get_state(0)
    plain.close()
    assert rng == rngp and t == tp and np.array_equal(y, yp)
