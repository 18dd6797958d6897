"""YAML front end: xmd's parameter files -> xmd_params.

Mirrors the reference's configuration path:
  - defaults merged with a user file: xmd/src/utils/yaml.cpp:4-42
    (`merge_yaml`), xmd/src/freeform/main.cpp:97-106;
  - "value unit-expr" quantities: xmd/src/utils/units.cpp:12-51
    (`quantity::quantity`), unit table xmd/include/xmd/utils/units.h:9-62;
  - section keys: xmd/src/freeform/params.cpp:7-309;
  - aa data: xmd/src/types/amino_acid.cpp:68-150;
  - lj force variants: xmd/src/forces/primitives/lj_variants.cpp:8-63.
The unit expressions are evaluated by a small recursive-descent parser of
this module (the reference uses the cparse library); no Python eval.
"""
import copy
import csv
import io
import math
import os
import re

import numpy as np
import yaml

from . import _abi

# amino-acid codes in aa_code order (xmd/include/xmd/types/amino_acid.h:10-13)
AA_NAMES = ["ALA", "ARG", "ASN", "ASP", "CYS", "GLU", "GLN", "GLY", "HIS", "ILE",
            "LEU", "LYS", "MET", "PHE", "PRO", "SER", "THR", "TRP", "TYR", "VAL"]
AA_INDEX = {n: k for k, n in enumerate(AA_NAMES)}
AA_LETTERS = "ARNDCEQGHILKMFPSTWYV"
POLARIZATION = {"polar": 1, "hydrophobic": 2}


def _unit_table():
    """Internal unit system: length 5 A, time tau, energy eps, kB = 1."""
    u = {}
    u["f77unit"] = 1.0
    u["A"] = u["f77unit"] / 5.0
    u["nm"] = 10.0 * u["A"]
    u["m"] = u["nm"] * 1.0e9
    u["ns"] = 1.0
    u["tau"] = 1.0 * u["ns"]
    u["micros"] = u["ns"] * 1.0e3
    u["ms"] = u["ns"] * 1.0e6
    u["s"] = u["ns"] * 1.0e9
    u["atom"] = 1.0
    u["mol"] = 6.02214076e23 / u["atom"]
    u["eps"] = 1.0
    u["kcal"] = u["eps"] * u["mol"] / 1.57824959
    u["J"] = u["kcal"] / 4184.0
    eps_kB = 1.0
    u["kB"] = u["eps"] / eps_kB
    u["K"] = 1.380649e-23 * u["J"] / u["kB"]
    u["kg"] = u["J"] * u["s"] * u["s"] / (u["m"] * u["m"])
    u["amu"] = u["kg"] * 0.99999999965e-3 / u["mol"]
    u["f77mass"] = u["eps"] * u["tau"] * u["tau"] / (u["f77unit"] * u["f77unit"])
    u["e"] = 1.0
    u["C"] = u["e"] / 1.602176634e-19
    u["Amp"] = u["C"] / u["s"]
    u["c"] = 299792458.0 * u["m"] / u["s"]
    u["H"] = u["kg"] * u["m"] * u["m"] / (u["s"] * u["s"] * u["Amp"] * u["Amp"])
    u["mu_0"] = 1.25663706212e-6 * u["H"] / u["m"]
    u["eps_0"] = 1.0 / (u["mu_0"] * u["c"] * u["c"])
    u["rad"] = 1.0
    u["deg"] = (2.0 * math.pi / 360.0) * u["rad"]
    return u


UNITS = _unit_table()
ANGSTROM = UNITS["A"]

_TOKEN = re.compile(r"\s*(?:(\d+\.?\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?)|([A-Za-z_][A-Za-z_0-9]*)|(\*\*|[-+*/()^]))")


def _tokenize(s):
    pos, out = 0, []
    s = s.strip()
    while pos < len(s):
        m = _TOKEN.match(s, pos)
        if not m or m.end() == pos:
            raise ValueError(f"cannot parse unit expression {s!r} at {pos}")
        num, ident, op = m.groups()
        if num is not None:
            out.append(("n", float(num)))
        elif ident is not None:
            out.append(("i", ident))
        else:
            out.append(("o", "**" if op == "^" else op))
        pos = m.end()
    return out


def evaluate(expr):
    """Evaluate an arithmetic expression over numbers and unit names."""
    toks = _tokenize(expr)
    k = [0]

    def peek():
        return toks[k[0]] if k[0] < len(toks) else (None, None)

    def take():
        t = peek()
        k[0] += 1
        return t

    def primary():
        kind, v = take()
        if kind == "n":
            return v
        if kind == "i":
            if v not in UNITS:
                raise ValueError(f"unknown unit {v!r} in {expr!r}")
            return UNITS[v]
        if (kind, v) == ("o", "("):
            x = add()
            if take() != ("o", ")"):
                raise ValueError(f"missing ')' in {expr!r}")
            return x
        raise ValueError(f"unexpected token {v!r} in {expr!r}")

    def unary():
        if peek() == ("o", "-"):
            take()
            return -unary()
        if peek() == ("o", "+"):
            take()
            return unary()
        return power()

    def power():
        base = primary()
        if peek() == ("o", "**"):
            take()
            return base ** unary()
        return base

    def mul():
        x = unary()
        while peek() in (("o", "*"), ("o", "/")):
            op = take()[1]
            y = unary()
            x = x * y if op == "*" else x / y
        return x

    def add():
        x = mul()
        while peek() in (("o", "+"), ("o", "-")):
            op = take()[1]
            y = mul()
            x = x + y if op == "+" else x - y
        return x

    val = add()
    if k[0] != len(toks):
        raise ValueError(f"trailing tokens in {expr!r}")
    return val


def quantity(value, default_unit=1.0):
    """`quantity(string, def_unit)`: magnitude, then optional unit after the
    first space (xmd/src/utils/units.cpp:30-49)."""
    if isinstance(value, (int, float)) and not isinstance(value, bool):
        return float(value) * default_unit
    s = str(value)
    if " " not in s:
        return evaluate(s) * default_unit
    mag, unit = s.split(" ", 1)
    return evaluate(mag) * evaluate(unit)


def merge_yaml(defaults, overrides):
    """Deep merge of mappings; overrides win (xmd/src/utils/yaml.cpp:4-42)."""
    if not isinstance(overrides, dict):
        return defaults if overrides is None else overrides
    if not isinstance(defaults, dict):
        return overrides
    out = {}
    for key, val in defaults.items():
        out[key] = merge_yaml(val, overrides[key]) if key in overrides else copy.deepcopy(val)
    for key, val in overrides.items():
        if key not in out:
            out[key] = copy.deepcopy(val)
    return out


def default_yaml():
    with open(os.path.join(os.path.dirname(__file__), "defaults.yml")) as f:
        return yaml.safe_load(f)


def load(path_or_dict=None):
    """Merged parameter tree (defaults + user file or dict)."""
    base = default_yaml()
    if path_or_dict is None:
        return base, "."
    if isinstance(path_or_dict, dict):
        return merge_yaml(base, path_or_dict), "."
    with open(path_or_dict) as f:
        user = yaml.safe_load(f) or {}
    return merge_yaml(base, user), os.path.dirname(os.path.abspath(path_or_dict))


def _csv_records(node, base_dir):
    text = str(node)
    if "\n" not in text and "," not in text:
        with open(os.path.join(base_dir, text)) as f:
            text = f.read()
    return list(csv.DictReader(io.StringIO(text)))


def _csv_rows(node, base_dir):
    """Raw comma-split rows after the header (the reference's readCSVRow,
    xmd/src/files/csv.cpp:28-69, without quoting), for tables read by
    position."""
    text = str(node)
    if "\n" not in text and "," not in text:
        with open(os.path.join(base_dir, text)) as f:
            text = f.read()
    lines = [ln for ln in text.splitlines() if ln.strip()]
    return [ln.split(",") for ln in lines[1:]]


def _heur_type(c1, c2):
    """heur_coeff_pair_type (params.cpp:58-67): G 0, P 1, anything else 2,
    from the first character of each type field."""
    cls = lambda f: {"G": 0, "P": 1}.get(f.strip()[:1], 2)
    return 3 * cls(c1) + cls(c2)


def heuristic_tables(tree, base_dir="."):
    """Coefficient tables of "heurestic angles" / "heurestic dihedrals"
    (params.cpp:69-124) by pair type, read by position (decision D29):
    7 polynomial coefficients a0..a6 per angle row (eps/rad^d), 6
    dihedral coefficients const, sin, cos, sin2, cos2, sin_cos (eps)."""
    ang = np.zeros((9, 7))
    dih = np.zeros((9, 6))
    for sec, table, ncoef, unit in (("heurestic angles", ang, 7, 1.0), ("heurestic dihedrals", dih, 6, UNITS["eps"])):
        node = (tree.get(sec) or {}).get("coefficients")
        if node is None:
            continue
        for row in _csv_rows(node, base_dir):
            vals = [f for f in row[2:] if f.strip()]
            if len(vals) < ncoef:
                raise ValueError(f"{sec}: row {','.join(row)} has fewer than {ncoef} coefficients")
            table[_heur_type(row[0], row[1])] = [quantity(v.strip(), unit) for v in vals[:ncoef]]
    return ang, dih


def _bool(x):
    if isinstance(x, bool):
        return x
    return str(x).strip().lower() in ("true", "yes", "on", "1")


def to_xmd_params(tree, base_dir=".", precision=_abi.PREC_FP64, rng_exact=True,
                  max_cutoff_override=0.0):
    """Fill an XmdParams from a merged parameter tree."""
    p = _abi.XmdParams()
    q = quantity

    lang = tree["langevin"]
    p.langevin_enabled = int(_bool(lang["enabled"]) and _bool(tree["integrator"]["enabled"]))
    p.gamma_factor = q(lang["gamma factor"])
    p.dt = q(tree["integrator"]["dt"])
    p.precision = int(precision)
    p.rng_exact = int(bool(rng_exact))
    p.pad_factor = q(tree["neighbor list"]["pad factor"])
    p.max_cutoff_override = float(max_cutoff_override)

    t = tree["tether forces"]
    p.tether_enabled = int(_bool(t["enabled"]))
    p.tether_H1, p.tether_H2 = q(t["H1"]), q(t["H2"])
    p.tether_def_length = q(t["default bond length"])
    c = tree["chirality"]
    p.chir_enabled, p.chir_e = int(_bool(c["enabled"])), q(c["e_chi"])
    a = tree["native angles"]
    p.nat_ang_enabled, p.nat_ang_k = int(_bool(a["enabled"])), q(a["k"])
    d = tree["complex native dihedrals"]
    p.cnd_enabled, p.cnd_CDA, p.cnd_CDB = int(_bool(d["enabled"])), q(d["CDA"]), q(d["CDB"])
    d = tree["simple native dihedrals"]
    p.snd_enabled, p.snd_CDH = int(_bool(d["enabled"])), q(d["CDH"])
    pa = tree["Pauli exclusion"]
    p.pauli_enabled, p.pauli_r_excl, p.pauli_depth = int(_bool(pa["enabled"])), q(pa["r_excl"]), q(pa["depth"])
    g = tree["native contacts"]
    p.go_enabled, p.go_depth = int(_bool(g["enabled"])), q(g["lj depth"])
    ss = tree["native ssbonds"]
    p.nat_ss_enabled, p.nat_ss_H1 = int(_bool(ss["enabled"])), q(ss["H1"])
    p.nat_ss_equil_dist = q(ss["equilibrium dist"])
    # params.cpp:223-270 (planes and tips go to the topology, walls_and_tips)
    w = tree["solid walls"]
    p.solid_walls_enabled = int(_bool(w["enabled"]))
    p.solid_walls_depth, p.solid_walls_cutoff = q(w["depth"]), q(w["cutoff"])
    w = tree["LJ attractive walls"]
    p.lj_walls_enabled = int(_bool(w["enabled"]))
    p.lj_walls_min_dist, p.lj_walls_breaking_factor = q(w["wall min dist"]), float(w["breaking factor"])
    p.lj_walls_phase_duration = q(w["phase duration"])
    w = tree["velocity AFM"]
    p.vel_afm_enabled, p.vel_afm_H1, p.vel_afm_H2 = int(_bool(w["enabled"])), q(w["H1"]), q(w["H2"])
    p.force_afm_enabled = int(_bool(tree["force AFM"]["enabled"]))
    p.heur_ang_enabled = int(_bool((tree.get("heurestic angles") or {}).get("enabled", False)))
    p.heur_dih_enabled = int(_bool((tree.get("heurestic dihedrals") or {}).get("enabled", False)))
    ang, dih = heuristic_tables(tree, base_dir)
    for k in range(9):
        for d in range(7):
            p.heur_ang_coeffs[k][d] = ang[k, d]
        for d in range(6):
            p.heur_dih_coeffs[k][d] = dih[k, d]
    e = tree["const ES"]
    p.const_es_enabled = int(_bool(e["enabled"]))
    p.const_es_screen, p.const_es_permittivity = q(e["screening distance"]), q(e["permittivity"])
    e = tree["relative ES"]
    p.rel_es_enabled = int(_bool(e["enabled"]))
    p.rel_es_screen, p.rel_es_factor = q(e["screening distance"]), q(e["factor"])

    aa = tree["amino acid data"]["amino acids"]
    for name, k in AA_INDEX.items():
        ent = aa.get(name, {}) or {}
        p.aa_charge[k] = q(ent["charge"], UNITS["e"]) if "charge" in ent else 0.0
        p.aa_polarization[k] = POLARIZATION.get(str(ent.get("polarization", "")), 0)
        lim = ent.get("contact limits", {}) or {}
        p.aa_limits[k][0] = int(lim.get("back", 0))
        p.aa_limits[k][1] = int(lim.get("side (all)", 0))
        p.aa_limits[k][2] = int(lim.get("side (polar)", 0))
        p.aa_limits[k][3] = int(lim.get("side (hydrophobic)", 0))

    lv = tree["lj force variants"]
    p.lj_bb_rmin, p.lj_bb_depth = q(lv["bb"]["r_min"]), q(lv["bb"]["depth"])
    p.lj_bs_rmin, p.lj_bs_depth = q(lv["bs"]["r_min"]), q(lv["bs"]["depth"])
    ss = lv["ss"]
    dflt = ss.get("default", {}) or {}
    rmax = [0.0] * _abi.NUM_SS
    depth = [0.0] * _abi.NUM_SS
    if "r_max" in dflt:
        rmax = [q(dflt["r_max"])] * _abi.NUM_SS
    if "depth" in dflt:
        depth = [q(dflt["depth"])] * _abi.NUM_SS
    per_pair = ss.get("per pair", {}) or {}
    for key, arr, unit in (("r_max", rmax, ANGSTROM), ("depth", depth, UNITS["eps"])):
        if key in per_pair:
            for rec in _csv_records(per_pair[key], base_dir):
                a1 = AA_INDEX[rec["type"].strip()]
                for a2n, a2 in AA_INDEX.items():
                    arr[_abi.NUM_AA * a1 + a2] = quantity(rec[a2n].strip(), unit)
    for k in range(_abi.NUM_SS):
        p.ss_rmax[k], p.ss_depth[k] = rmax[k], depth[k]
    p.ss_sinking = int(_bool(ss["use sinking variant"]))

    pid = tree["pseudo-improper dihedral potential"]
    p.pid_enabled = int(_bool(pid["enabled"]))
    p.pid_lambda_version = _abi.LAMBDA_ALGEBRAIC if str(pid.get("lambda version", "")).strip() == "algebraic" \
        else _abi.LAMBDA_COSINE
    for tag, key in (("bbp", "bb+"), ("bbm", "bb-")):
        sec = pid[key]
        setattr(p, f"pid_{tag}_alpha", q(sec["alpha"]))
        setattr(p, f"pid_{tag}_psi0", q(sec["psi_0"]))
        setattr(p, f"pid_{tag}_rmin", q(sec["r_min"]))
        setattr(p, f"pid_{tag}_depth", q(sec["depth"]))
    p.pid_ss_alpha, p.pid_ss_psi0 = q(pid["ss"]["alpha"]), q(pid["ss"]["psi_0"])

    qa = tree["quasi-adiabatic potential"]
    p.qa_enabled = int(_bool(qa["enabled"]))
    p.qa_phase_duration = q(qa["phase duration"])
    p.qa_breaking_factor = float(qa["breaking factor"])
    p.qa_min_cos_hr = float(qa["min |cos(h, r)|"])
    p.qa_min_cos_hh = float(qa["min |cos(h, h)| for bb"])
    p.qa_max_cos_nr = float(qa["max cos(n, r)"])
    return p


def _vec3(node):
    """A YAML [x, y, z] of quantities (xmd/src/types/vec3.cpp:10-19)."""
    return [quantity(v) for v in node]


def walls_and_tips(tree):
    """Wall planes and AFM tips of a parameter tree, in the topology's
    layout: planes (k, 6) = origin, normal; velocity tips (residue index,
    origin or NaN, velocity); force tips (residue index, force)."""
    def planes(sec):
        return np.array([_vec3(pl["origin"]) + _vec3(pl["normal"]) for pl in (sec.get("planes") or [])],
                        dtype=np.float64).reshape(-1, 6)
    vt = tree["velocity AFM"].get("AFM tips") or []
    ft = tree["force AFM"].get("AFM tips") or []
    return dict(
        solid_walls=planes(tree["solid walls"]), lj_walls=planes(tree["LJ attractive walls"]),
        vel_afm_idx=np.array([int(t["pulled idx"]) for t in vt], dtype=np.int32),
        vel_afm_orig=np.array([_vec3(t["origin"]) if t.get("origin") is not None else [np.nan] * 3 for t in vt],
                              dtype=np.float64).reshape(-1, 3),
        vel_afm_v=np.array([_vec3(t["v"]) for t in vt], dtype=np.float64).reshape(-1, 3),
        force_afm_idx=np.array([int(t["pulled idx"]) for t in ft], dtype=np.int32),
        force_afm_F=np.array([_vec3(t["F"]) for t in ft], dtype=np.float64).reshape(-1, 3),
    )


def from_yaml(path_or_dict=None, **kw):
    tree, base = load(path_or_dict)
    return to_xmd_params(tree, base, **kw), tree
