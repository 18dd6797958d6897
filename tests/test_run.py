"""The run driver (xmd_amd/run.py) on the CPU with a stand-in engine that
advances time as the engine does (t += dt per step, FP64): reports at the
reference's times (t - last_t >= period from the first step, main.cpp:
1051-1055 with report_stats.cpp / export_pdb.cpp / report_structure.cpp),
the run ending at the first step with t >= total time, files per replica."""
import math
import os

import numpy as np
import pytest

from xmd_amd import run

MODELS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models")


class FakeEngine:
    def __init__(self, params, system):
        self.dt, self.t, self.system = params.dt, 0.0, system
        self.steps, self.evals = 0, 0

    def step(self, n):
        for _ in range(n):
            self.t += self.dt
        self.steps += n

    def get_state(self, k):
        sl = self.system.replica_slice(k)
        y = np.zeros((6, 3, sl.stop - sl.start))
        y[0] = self.system.r0[:, sl] + self.t
        return y, None, 0, self.t

    def eval_forces(self):
        self.evals += 1

    def observables(self, thr):
        rows = np.zeros((self.system.n_replicas, 8))
        rows[:, 0] = self.t
        rows[:, 1] = -self.t
        rows[:, 4] = 1.0
        return rows

    def close(self):
        pass


def expected_times(total, dt, period):
    t, last, out = 0.0, -math.inf, []
    while t < total:
        if t - last >= period:
            out.append(t)
            last = t
        t += dt
    return out, t


@pytest.mark.parametrize("replicas", [1, 2])
def test_report_schedule(tmp_path, replicas):
    inp = {"model": {"seqfile": "glut/glut.yml"}, "general": {"total time": "1 tau"},
           "report stats": {"enabled": True, "path": "out.csv", "exec period": "0.2 tau"},
           "report structure": {"enabled": True, "path format": "structure/%.3f.yml", "exec period": "0.45 tau"},
           "export pdb": {"enabled": True, "path": "traj.pdb", "exec period": "0.3 tau"}}
    eng = run.simulate(inp, replicas=replicas, base_dir=MODELS, out_dir=str(tmp_path), engine_cls=FakeEngine)
    dt = eng.dt
    st, t_end = expected_times(1.0, dt, 0.2)
    assert eng.t == t_end and eng.steps == int(round(t_end / dt))
    for k in range(replicas):
        suffix = "" if k == 0 else f".{k}"
        rows = open(tmp_path / f"out{suffix}.csv").read().splitlines()
        assert rows[0] == "t[tau],V[eps],K[eps],E[eps],Rg[A],W[A]"
        assert [r.split(",")[0] for r in rows[1:]] == ["%f" % t for t in st]
        assert [r.split(",")[1] for r in rows[1:]] == ["%f" % -t for t in st]
        ss, _ = expected_times(1.0, dt, 0.45)
        files = sorted(os.listdir(tmp_path / "structure"))
        assert [f for f in files if f.endswith(f"{suffix}.yml") and (k or f.count(".") == 2)] == \
            sorted(("%.3f" % t) + f"{suffix}.yml" for t in ss)
        ps, _ = expected_times(1.0, dt, 0.3)
        pdb = open(tmp_path / f"traj{suffix}.pdb").read()
        assert pdb.count("ENDMDL") == len(ps)
    # forces evaluated once per reporting point, not per step
    assert eng.evals == len(sorted(set(st) | set(ss) | set(ps)))
