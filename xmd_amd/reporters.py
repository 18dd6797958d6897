"""File reporters of a run: the stats CSV, the structure YAML and the PDB
trajectory, in the reference's formats, fed by the engine's state and its
GPU observables (xmd_observables).

  report_stats      xmd/src/io/report_stats.cpp:8-41
  report_structure  xmd/src/io/report_structure.cpp:21-215 (main.cpp:330-343
                    wires it without the sync and QA sections; both are
                    optional arguments here)
  export_pdb        xmd/src/io/export_pdb.cpp:10-34, pdb.cpp:110-168,474-506
                    and the record writers of records.cpp / fields.cpp

Numbers are printed as the reference prints them: std::to_string (six
decimals) for CSV fields, printf widths for PDB columns.  Each reporter
writes when t - last_t >= period, starting from the first call.
"""
import math
import os

import numpy as np

from . import params as P
from .model import UNKNOWN, BACK_BACK, BACK_SIDE, SIDE_BACK, SIDE_SIDE, NAT_SS

_A = P.ANGSTROM
_TYPE_NAME = {UNKNOWN: "", BACK_BACK: "BACK_BACK", BACK_SIDE: "BACK_SIDE", SIDE_BACK: "SIDE_BACK",
              SIDE_SIDE: "SIDE_SIDE", NAT_SS: "NAT_SS"}


def _f(x):
    """std::to_string(double)."""
    return "%f" % x


class _Periodic:
    def __init__(self, period):
        self.period = period
        self.last_t = -math.inf

    def due(self, t):
        return t - self.last_t >= self.period


# ----------------------------------------------------------------- stats ---
STATS_HEADER = ["t[tau]", "V[eps]", "K[eps]", "E[eps]", "Rg[A]", "W[A]"]


def stats_row(obs):
    """One CSV row from a row of xmd_observables (t, V, K, E, Rg, W, ...).
    W is divided by one angstrom as the reference does (its unit is A^2)."""
    t, V, K, E, Rg, W = obs[:6]
    return [_f(t), _f(V), _f(K), _f(E), _f(Rg / _A), _f(W / _A)]


class StatsReport(_Periodic):
    """report_stats: a CSV with a header line, one row per report."""

    def __init__(self, path, period):
        super().__init__(period)
        self.path = path
        self.first = True

    def __call__(self, obs):
        t = float(obs[0])
        if not self.due(t):
            return False
        with open(self.path, "w" if self.first else "a") as f:
            if self.first:
                f.write(",".join(STATS_HEADER) + "\n")
            f.write(",".join(stats_row(obs)) + "\n")
        self.first = False
        self.last_t = t
        return True


# ------------------------------------------------------------- structure ---
def _csv_block(header, rows):
    out = "|-\n  " + ",".join(header)
    for r in rows:
        out += "\n  " + ",".join(r)
    return out


_SYNC_FIELDS = ["back", "side-all", "side-polar", "side-hydrophobic"]
_QA_TYPE = {1: "BACK_BACK", 2: "BACK_SIDE", 3: "SIDE_BACK"}  # others SIDE_SIDE (contact_type.h:14-30)
_QA_STATUS = {0: "FORMING_OR_FORMED", 1: "BREAKING"}


def structure_yaml(m, r, active_threshold, qa_contacts=None, sync=None):
    """The structure report of model m at positions r ((3, n), model residue
    order): residues, native contacts (current distance without periodic
    images, active below thr 2^(-1/6) d_nat), angles, dihedrals.  The
    reporter's optional parts (report_structure.cpp:28-51,92-141), which
    main.cpp leaves unset: `sync` ((n, 4) budgets, as from get_qa_state)
    adds the four budget columns to the residues; `qa_contacts` (the
    contact records of get_qa_state) adds the "QA contacts" section.  Its
    idx is the contact's place among the live contacts (the engine keeps
    its contact set compact, so the reference's vacant slots have no
    counterpart)."""
    header = ["idx", "x[A]", "y[A]", "z[A]"] + (_SYNC_FIELDS if sync is not None else [])
    res = []
    for i in range(r.shape[1]):
        row = [str(i), _f(r[0, i] / _A), _f(r[1, i] / _A), _f(r[2, i] / _A)]
        if sync is not None:
            row += [str(int(v)) for v in np.asarray(sync).reshape(-1, 4)[i]]
        res.append(row)
    out = "residues: " + _csv_block(header, res)
    cut = active_threshold * 2.0 ** (-1.0 / 6.0)
    rows = []
    for k, (i1, i2, length, typ) in enumerate(m.contacts):
        d = float(np.linalg.norm(r[:, i1] - r[:, i2]))
        rows.append([str(k), str(i1), str(i2), _TYPE_NAME[typ], _f(length / _A), _f(d / _A),
                     "1" if d < cut * length else "0"])
    out += "\n\nnative contacts: " + _csv_block(
        ["idx", "i1", "i2", "type", "nat_dist[A]", "cur_dist[A]", "active"], rows)
    if qa_contacts is not None:
        rows = [[str(k), str(int(c["i1"])), str(int(c["i2"])), _QA_TYPE.get(int(c["type"]), "SIDE_SIDE"),
                 _f(float(c["ref_time"])), _QA_STATUS[int(c["status"])]] for k, c in enumerate(qa_contacts)]
        out += "\n\nQA contacts: " + _csv_block(["idx", "i1", "i2", "type", "t0[tau]", "status"], rows)
    rows = []
    for k, (i1, i2, i3, th) in enumerate(m.angles):
        r12, r23 = r[:, i2] - r[:, i1], r[:, i3] - r[:, i2]
        c = -float(r12 @ r23) / (np.linalg.norm(r12) * np.linalg.norm(r23))
        rows.append([str(k), str(i1), str(i2), str(i3), "" if th is None else _f(th), _f(math.acos(c))])
    out += "\n\nangles: " + _csv_block(["idx", "i1", "i2", "i3", "nat_theta[rad]", "cur_theta[rad]"], rows)
    rows = []
    for k, (i1, i2, i3, i4, ph) in enumerate(m.dihedrals):
        r12, r23, r34 = r[:, i2] - r[:, i1], r[:, i3] - r[:, i2], r[:, i4] - r[:, i3]
        x1, x2 = np.cross(r12, r23), np.cross(r23, r34)
        c = float((x1 / np.linalg.norm(x1)) @ (x2 / np.linalg.norm(x2)))
        phi = math.acos(c) if -1.0 <= c <= 1.0 else math.nan
        if float(x1 @ r34) < 0.0:
            phi = -phi
        rows.append([str(k), str(i1), str(i2), str(i3), str(i4), "" if ph is None else _f(ph), _f(phi)])
    out += "\n\ndihedrals: " + _csv_block(["idx", "i1", "i2", "i3", "i4", "nat_phi[rad]", "cur_phi[rad]"], rows)
    return out


class StructureReport(_Periodic):
    """report_structure: one YAML file per report, named path_fmt % t."""

    def __init__(self, path_fmt, period, active_threshold=1.5):
        super().__init__(period)
        self.path_fmt = path_fmt
        self.thr = active_threshold

    def __call__(self, t, m, r, qa_contacts=None, sync=None):
        if not self.due(t):
            return None
        path = self.path_fmt % t
        if os.path.dirname(path):
            os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            f.write(structure_yaml(m, r, self.thr, qa_contacts, sync))
        self.last_t = t
        return path


# ------------------------------------------------------------------- PDB ---
def _put(line, i, j, text):
    """Write text right-aligned into columns i..j (1-based), as the
    reference's "%*s" field writers do."""
    s = "%*s" % (j - i + 1, text)
    return line[:i - 1] + s + line[i - 1 + len(s):]


def pdb_model_text(m, r, serial):
    """One MODEL of the PDB trajectory (pdb.cpp:110-168,474-506): chains
    A, B, ... in model order, residues numbered 1.. over the whole model,
    one CA atom each (serials 1..), a TER per chain (serial 0, as the
    reference's chain record leaves it), ENDMDL; no trailing newline."""
    out = _put(" " * 80, 1, 6, "MODEL ")
    out = _put(out, 11, 14, "%d" % serial)
    num = 1
    for c, chain in enumerate(m.chains):
        cid = chr(ord("A") + c)
        for res in chain:
            name = P.AA_NAMES[m.atype[res]]
            line = _put(" " * 80, 1, 6, "ATOM  ")
            line = _put(line, 7, 11, "%d" % num)
            line = _put(line, 13, 16, "CA")
            line = _put(line, 18, 20, name)
            line = line[:21] + cid + line[22:]
            line = _put(line, 23, 26, "%d" % num)
            for a, (i, j) in enumerate(((31, 38), (39, 46), (47, 54))):
                line = _put(line, i, j, "%8.3f" % (r[a, res] / _A))
            out += "\n" + line
            last_name, last_num = name, num
            num += 1
        line = _put(" " * 80, 1, 6, "TER   ")
        line = _put(line, 7, 11, "0")
        line = _put(line, 18, 20, last_name)
        line = line[:21] + cid + line[22:]
        line = _put(line, 23, 26, "%d" % last_num)
        out += "\n" + line
    out += "\n" + _put(" " * 80, 1, 6, "ENDMDL")
    return out


class PdbExport(_Periodic):
    """export_pdb: MODEL records appended to one file, serials from 1."""

    def __init__(self, path, period):
        super().__init__(period)
        self.path = path
        self.serial = 1

    def __call__(self, t, m, r):
        if not self.due(t):
            return False
        with open(self.path, "w" if self.serial == 1 else "a") as f:
            if self.serial > 1:
                f.write("\n")
            f.write(pdb_model_text(m, r, self.serial))
        self.serial += 1
        self.last_t = t
        return True


def model_positions(engine, system, rep):
    """Positions of replica rep in model residue order ((3, n), internal
    units): the topology keeps chains in model order, so this is the
    replica's y0."""
    y, _, _, _ = engine.get_state(rep)
    return y[0]
