"""A whole run of an xmd input file on the engine: the loop of the
reference's freeform main (xmd/src/freeform/main.cpp:1028-1076) with its
three file reporters (export pdb, report stats, report structure) at their
periods, until `general.total time`.

At every reporting point the reference has just evaluated the forces at the
current (predicted) positions; its reporters then write t, V of those
forces and the positions (main.cpp:1051-1055).  Here the engine steps in
chunks up to the next reporting time, and at a reporting point the forces
are evaluated at the current state (xmd_eval_forces, which changes no
state) so that V and the positions belong to the same point, as in the
reference.  The reporters' `t - last_t >= period` test uses the engine's
FP64 time.  Paths are taken relative to the input file's directory (the
reference resolves them against the YAML file, params.cpp:199-221), or
`out_dir`; replica k > 0 of a batch writes to the same names with `.k`
before the extension.
"""
import math
import os

from . import model as M, params as P, reporters as R


def _replica_path(path, k):
    if k == 0:
        return path
    stem, ext = os.path.splitext(path)
    return f"{stem}.{k}{ext}"


def simulate(input_file, replicas=1, base_dir=None, out_dir=None, total_time=None, engine_cls=None, **kw):
    """Run the input file (a path or a dict merged over the defaults) on the
    engine and write its reports.  Model paths are taken from base_dir
    (default: the input file's directory), reports go to out_dir (default:
    base_dir).  Returns the engine, still open, after the last step
    (callers close it)."""
    tree, base = P.load(input_file)
    base = base if base_dir is None else base_dir
    out_dir = base if out_dir is None else out_dir
    params, system, models = M.from_input(input_file, base_dir=base, replicas=replicas, **kw)
    if engine_cls is None:
        from .engine import Engine as engine_cls
    eng = engine_cls(params, system)
    total = P.quantity(tree["general"]["total time"]) if total_time is None else float(total_time)
    dt = params.dt
    thr = float(tree["native contacts"].get("active threshold", 1.5))

    reps = []  # (kind, per-replica reporter list)
    sec = tree.get("report stats") or {}
    if P._bool(sec.get("enabled", False)):
        reps.append(("stats", [R.StatsReport(os.path.join(out_dir, _replica_path(str(sec["path"]), k)),
                                             P.quantity(sec["exec period"])) for k in range(replicas)]))
    sec = tree.get("report structure") or {}
    if P._bool(sec.get("enabled", False)):
        reps.append(("structure", [R.StructureReport(os.path.join(out_dir, _replica_path(str(sec["path format"]), k)),
                                                     P.quantity(sec["exec period"]), thr) for k in range(replicas)]))
    sec = tree.get("export pdb") or {}
    if P._bool(sec.get("enabled", False)):
        reps.append(("pdb", [R.PdbExport(os.path.join(out_dir, _replica_path(str(sec["path"]), k)),
                                         P.quantity(sec["exec period"])) for k in range(replicas)]))

    def now():
        return eng.get_state(0)[3]

    t = now()
    while t < total:
        if any(r[0].due(t) for _, r in reps):
            eng.eval_forces()
            obs = eng.observables(thr)
            for k in range(replicas):
                r = eng.get_state(k)[0][0]  # y0: positions in model residue order
                for kind, lst in reps:
                    if kind == "stats":
                        lst[k](obs[k])
                    else:
                        lst[k](float(obs[k][0]), models[k], r)
        # steps to the earlier of the next reporting time and the end,
        # stopping one step short of a reporting time so that it is met
        # exactly by the single steps that follow
        horizon = min([lst[0].last_t + lst[0].period for _, lst in reps] + [total])
        n = max(1, int(math.ceil((horizon - t) / dt)) - 1)
        eng.step(n)
        t = now()
    return eng
