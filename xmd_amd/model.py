"""Protein models and the engine topology built from them.

Host-side front end of the engine (file formats either side of the step
path), following the reference's model pipeline:

  sequence file + contact maps   xmd/src/files/seq_file.cpp:9-134,
                                 xmd/src/files/map_file.cpp:68-112
  PDB file -> model              xmd/src/files/pdb.cpp:11-108,198-323,
                                 xmd/src/files/records.cpp, fields.cpp
  morph into a random chain      xmd/src/model/model.cpp:92-174
  model -> per-term inputs       xmd/src/freeform/main.cpp:113-206,372-553

Model lengths are in internal units (5 A), angles in radians.  Where the
reference's behaviour is undefined or self-inconsistent the choice made here
is stated at the function (the same list is kept with the engine's other
reference decisions).
"""
import csv
import io
import math
import os
import re
from dataclasses import dataclass, field

import numpy as np
import yaml

from . import params as P

# contact types of xmd::model (xmd/include/xmd/model/model.h)
UNKNOWN, BACK_BACK, BACK_SIDE, SIDE_BACK, SIDE_SIDE, NAT_SS = range(6)
BACKBONE_ATOMS = ("N", "CA", "C", "O", "OXT")  # pdb.cpp:551-557


@dataclass
class Model:
    """Residues in chain order; index lists refer to model residue indices."""
    atype: list = field(default_factory=list)       # aa code per residue
    pos: list = field(default_factory=list)         # (x, y, z) per residue
    chains: list = field(default_factory=list)      # list of residue index lists
    tethers: list = field(default_factory=list)     # (i1, i2, length or None)
    contacts: list = field(default_factory=list)    # (i1, i2, length, type)
    angles: list = field(default_factory=list)      # (i1, i2, i3, theta or None)
    dihedrals: list = field(default_factory=list)   # (i1, i2, i3, i4, phi or None)
    box: tuple = (0.0, 0.0, 0.0)

    @property
    def n_residues(self):
        return len(self.atype)

    def add_residue(self, code, pos=(0.0, 0.0, 0.0)):
        self.atype.append(int(code))
        self.pos.append(tuple(float(x) for x in pos))
        return len(self.atype) - 1


# ------------------------------------------------------------- rand_gen ---
class RandGen:
    """The reference's generator (xmd/include/xmd/random/rand_gen.h:7-69):
    xorshift state update with a multiplied output, uniforms from the top 32
    output bits, Box-Muller normals in double, points on the unit sphere."""
    MASK = (1 << 64) - 1

    def __init__(self, seed):
        self.state = int(seed) & self.MASK

    def next(self):
        s = self.state
        res = (s * 0xD989BCACC137DCD5) & self.MASK
        s ^= s >> 11
        s ^= (s << 31) & self.MASK
        s ^= s >> 18
        self.state = s
        return res

    def uniform(self, a=None, b=None):
        u = float(self.next() >> 32) * (1.0 / 4294967296.0)
        return u if a is None else (b - a) * u + a

    def _radius_angle(self):
        r1, r2 = self.uniform(), self.uniform()
        r = math.sqrt(-2.0 * math.log(r1)) if r1 > 0.0 else math.inf
        return r, (2.0 * math.pi) * r2

    def normal(self):
        r, t = self._radius_angle()
        return r * math.cos(t)

    def normal_x2(self):
        r, t = self._radius_angle()
        return r * math.cos(t), r * math.sin(t)

    def sphere(self):
        x, y = self.normal_x2()
        z = self.normal()
        n = math.sqrt(x * x + y * y + z * z)
        return (x / n, y / n, z / n)


# -------------------------------------------------------- sequence files ---
def _csv_text(node, base_dir):
    """A CSV table given inline or as {from-file: path} (map_file.cpp:70-80;
    a relative path is taken from the map's directory)."""
    if node is None:
        return []
    if isinstance(node, dict) and "from-file" in node:
        with open(os.path.join(base_dir, node["from-file"])) as f:
            text = f.read()
    else:
        text = str(node)
    rows = csv.reader(io.StringIO(text.strip()))
    header = None
    out = []
    for row in rows:
        if not row or all(not c.strip() for c in row):
            continue
        row = [c.strip() for c in row]
        if header is None:
            header = row
            continue
        out.append(dict(zip(header, row)))
    return out


def read_map(node, base_dir="."):
    """Contacts, angles and dihedrals of a contact map (map_file.cpp:68-112).
    Numbers are taken as they stand, as the reference does: lengths in
    internal units, angles in radians.  A missing section is empty."""
    contacts = [(int(r["i1"]), int(r["i2"]), float(r["length"])) for r in _csv_text(node.get("contacts"), base_dir)]
    angles = [(int(r["i1"]), int(r["i2"]), int(r["i3"]), float(r["theta"]))
              for r in _csv_text(node.get("angles"), base_dir)]
    dihedrals = [(int(r["i1"]), int(r["i2"]), int(r["i3"]), int(r["i4"]), float(r["phi"]))
                 for r in _csv_text(node.get("dihedrals"), base_dir)]
    return contacts, angles, dihedrals


def read_seq_file(path):
    """Model of a sequence file (seq_file.cpp:9-134): chains of one-letter
    codes (whitespace removed), contact maps per chain (`from-file` relative
    to the sequence file, optional `shift` of the map's residue indices),
    every consecutive pair tethered with no native length, and the angles and
    dihedrals not given by a map added without a native value."""
    with open(path) as f:
        doc = yaml.safe_load(f)
    base = os.path.dirname(os.path.abspath(path))
    m = Model()
    for chain in doc["model"]["chains"]:
        seq = re.sub(r"\s+", "", str(chain["seq"]))
        idx = [m.add_residue(P.AA_LETTERS.index(c)) for c in seq]
        m.chains.append(idx)
        taken_ang, taken_dih = set(), set()
        for _, entry in (chain.get("maps") or {}).items():
            if "from-file" in entry:
                mpath = os.path.join(base, entry["from-file"])
                with open(mpath) as f:
                    mnode = yaml.safe_load(f)
                mbase = os.path.dirname(mpath)
            else:
                mnode, mbase = entry["map"], base
            cs, an, dh = read_map(mnode, mbase)
            sh = int(entry.get("shift", 0))
            for i1, i2, length in cs:
                m.contacts.append((idx[i1 + sh], idx[i2 + sh], length, UNKNOWN))
            for i1, i2, i3, th in an:
                key = (idx[i1 + sh], idx[i2 + sh], idx[i3 + sh])
                m.angles.append(key + (th,))
                taken_ang.add(key)
            for i1, i2, i3, i4, ph in dh:
                key = (idx[i1 + sh], idx[i2 + sh], idx[i3 + sh], idx[i4 + sh])
                m.dihedrals.append(key + (ph,))
                taken_dih.add(key)
        for k in range(len(idx) - 1):
            m.tethers.append((idx[k], idx[k + 1], None))
            if k + 2 < len(idx):
                key = (idx[k], idx[k + 1], idx[k + 2])
                if key not in taken_ang:
                    m.angles.append(key + (None,))
                if k + 3 < len(idx):
                    key = (idx[k], idx[k + 1], idx[k + 2], idx[k + 3])
                    if key not in taken_dih:
                        m.dihedrals.append(key + (None,))
    return m


# ------------------------------------------------------------ PDB files ---
def _field(line, i, j):
    """Columns i..j (1-based, inclusive) of a record line."""
    return line[i - 1:j]


def _int(line, i, j):
    return int(_field(line, i, j))


def _real(line, i, j):
    return float(_field(line, i, j))


@dataclass
class _PdbResidue:
    name: str
    atoms: list = field(default_factory=list)  # atom serials in record order


@dataclass
class _PdbChain:
    atoms: dict = field(default_factory=dict)     # serial -> (name, pos, res_seq)
    residues: dict = field(default_factory=dict)  # res_seq -> _PdbResidue


def read_pdb(path_or_text):
    """Model of a PDB file (pdb.cpp:11-108, to_model :198-323).

    Records (records.cpp) are recognised by columns 1-6 exactly as the
    reference does: ATOM, SSBOND (with CYS at 12-14), LINK, CRYST1, TER;
    END stops reading.  Residues take their CA atom; consecutive residues
    are tethered at their distance, the angle of every consecutive triple is
    acos(u12 . u23) (the angle between the bond vectors, as the reference
    computes it) and every consecutive quad gets its dihedral.  SSBOND pairs
    become NAT_SS contacts, LINK pairs contacts typed by whether the linked
    atoms are backbone atoms, both at the CA distance, each residue pair once.

    The reference keeps chains and residues in hash maps and walks them in
    hash order; here chains are taken in order of first appearance and the
    residues of a chain by ascending sequence number."""
    if "\n" in path_or_text:
        lines = path_or_text.splitlines()
    else:
        with open(path_or_text) as f:
            lines = f.read().splitlines()
    chains, order = {}, []
    ssbonds, links = {}, []
    cryst1 = (0.0, 0.0, 0.0)

    def chain_of(cid):
        if cid not in chains:
            chains[cid] = _PdbChain()
            order.append(cid)
        return chains[cid]

    for line in lines:
        rec = line[:6]
        if rec == "REMARK":
            continue
        if rec == "ATOM  ":
            serial = _int(line, 7, 11)
            name = _field(line, 13, 16).strip()
            res_name = _field(line, 18, 20).strip()
            cid = line[21]
            res_seq = _int(line, 23, 26)
            pos = (_real(line, 31, 38) * P.ANGSTROM, _real(line, 39, 46) * P.ANGSTROM,
                   _real(line, 47, 54) * P.ANGSTROM)
            c = chain_of(cid)
            res = c.residues.setdefault(res_seq, _PdbResidue(res_name))
            c.atoms[serial] = (name, pos, res_seq)
            res.atoms.append(serial)
        elif rec == "SSBOND":
            serial = _int(line, 8, 10)
            if _field(line, 12, 14) != "CYS":
                continue
            ssbonds[serial] = ((line[15], _int(line, 18, 21)), (line[29], _int(line, 32, 35)))
        elif rec == "LINK  ":
            links.append(((line[21], _int(line, 23, 26), _field(line, 13, 16).strip()),
                          (line[51], _int(line, 53, 56), _field(line, 43, 46).strip())))
        elif rec == "CRYST1":
            cryst1 = (_real(line, 7, 15) * P.ANGSTROM, _real(line, 16, 24) * P.ANGSTROM,
                      _real(line, 25, 33) * P.ANGSTROM)
        elif rec == "TER   ":
            _int(line, 7, 11)
        elif rec == "END   ":
            break

    m = Model(box=cryst1)
    res_index = {}
    for cid in order:
        c = chains[cid]
        idx = []
        for res_seq in sorted(c.residues):
            res = c.residues[res_seq]
            ca = next((c.atoms[s][1] for s in res.atoms if c.atoms[s][0] == "CA"), None)
            if ca is None:
                raise ValueError(f"residue {res.name} {res_seq} of chain {cid!r} has no CA atom")
            k = m.add_residue(P.AA_INDEX[res.name], ca)
            res_index[(cid, res_seq)] = k
            idx.append(k)
        m.chains.append(idx)
        p = [np.array(m.pos[k]) for k in idx]
        for a in range(len(idx) - 1):
            m.tethers.append((idx[a], idx[a + 1], float(np.linalg.norm(p[a + 1] - p[a]))))
            if a + 2 < len(idx):
                u12 = (p[a + 1] - p[a]) / np.linalg.norm(p[a + 1] - p[a])
                u23 = (p[a + 2] - p[a + 1]) / np.linalg.norm(p[a + 2] - p[a + 1])
                theta = math.acos(min(max(float(u12 @ u23), -1.0), 1.0))
                m.angles.append((idx[a], idx[a + 1], idx[a + 2], theta))
                if a + 3 < len(idx):
                    r12, r23, r34 = p[a + 1] - p[a], p[a + 2] - p[a + 1], p[a + 3] - p[a + 2]
                    x1, x2 = np.cross(r12, r23), np.cross(r23, r34)
                    cphi = float((x1 / np.linalg.norm(x1)) @ (x2 / np.linalg.norm(x2)))
                    phi = math.acos(min(max(cphi, -1.0), 1.0))
                    if float(x1 @ r34) < 0.0:
                        phi = -phi
                    m.dihedrals.append((idx[a], idx[a + 1], idx[a + 2], idx[a + 3], phi))

    def ca_dist(i, j):
        return float(np.linalg.norm(np.array(m.pos[i]) - np.array(m.pos[j])))

    seen = set()
    for serial in sorted(ssbonds):
        (c1, s1), (c2, s2) = ssbonds[serial]
        i, j = sorted((res_index[(c1, s1)], res_index[(c2, s2)]))
        if (i, j) in seen:
            continue
        seen.add((i, j))
        m.contacts.append((i, j, ca_dist(i, j), NAT_SS))
    for (c1, s1, a1), (c2, s2, a2) in links:
        i, j = res_index[(c1, s1)], res_index[(c2, s2)]
        b1, b2 = a1 in BACKBONE_ATOMS, a2 in BACKBONE_ATOMS
        if i > j:
            i, j, b1, b2 = j, i, b2, b1
        if (i, j) in seen:
            continue
        seen.add((i, j))
        t = BACK_BACK if b1 and b2 else BACK_SIDE if b1 else SIDE_BACK if b2 else SIDE_SIDE
        m.contacts.append((i, j, ca_dist(i, j), t))
    return m


# --------------------------------------------------------- morph to SAW ---
def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _norm(a):
    return math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])


def _quat(angle, axis):
    h = 0.5 * angle
    s = math.sin(h)
    return (math.cos(h), s * axis[0], s * axis[1], s * axis[2])


def _qmul(a, b):
    return (a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3],
            a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2],
            a[0] * b[2] + a[2] * b[0] + a[3] * b[1] - a[1] * b[3],
            a[0] * b[3] + a[3] * b[0] + a[1] * b[2] - a[2] * b[1])


def _qrot(q, v):
    u = q[1:]
    t = _cross(u, v)
    t = (2.0 * t[0], 2.0 * t[1], 2.0 * t[2])
    w = _cross(u, t)
    return (v[0] + q[0] * t[0] + w[0], v[1] + q[0] * t[1] + w[1], v[2] + q[0] * t[2] + w[2])


def morph_into_saw(m, gen, bond_length, residue_density, infer_box):
    """Lay every chain out as a random walk (model.cpp:92-174): the first
    residue uniform in a cube of side (N/density)^(1/3) about the origin,
    a direction uniform on the sphere, each next residue one bond along it
    (`bond_length`, or the model's own distance when None), the direction
    then tilted by U(-15, 15) degrees about a perpendicular and turned by
    U(-180, 180) degrees about itself.  The draws come from `gen` in the
    reference's order, so its state afterwards is the noise stream's start.
    With infer_box the box takes the extents the reference computes: every
    axis from the x coordinates, each maximum started at the smallest
    positive double (model.cpp:142-168)."""
    max_spread, max_around = math.pi / 12.0, math.pi
    cell_a = (m.n_residues / residue_density) ** (1.0 / 3.0)
    orig = [tuple(p) for p in m.pos]
    for chain in m.chains:
        pos = (gen.uniform(-cell_a / 2.0, cell_a / 2.0), gen.uniform(-cell_a / 2.0, cell_a / 2.0),
               gen.uniform(-cell_a / 2.0, cell_a / 2.0))
        d = gen.sphere()
        for k, res in enumerate(chain):
            nxt = pos
            if k + 1 < len(chain):
                if bond_length is not None:
                    bl = bond_length
                else:
                    a, b = orig[res], orig[chain[k + 1]]
                    bl = _norm((b[0] - a[0], b[1] - a[1], b[2] - a[2]))
                nxt = (pos[0] + d[0] * bl, pos[1] + d[1] * bl, pos[2] + d[2] * bl)
                spread = gen.uniform(-max_spread, max_spread)
                p1, p2 = _cross(d, (1.0, 0.0, 0.0)), _cross(d, (0.0, 1.0, 0.0))
                perp = p1 if _norm(p1) > _norm(p2) else p2
                n = _norm(perp)
                perp = (perp[0] / n, perp[1] / n, perp[2] / n)
                around = gen.uniform(-max_around, max_around)
                q = _qmul(_quat(around, d), _quat(spread, perp))
                d = _qrot(q, d)
                n = _norm(d)
                d = (d[0] / n, d[1] / n, d[2] / n)
            m.pos[res] = pos
            pos = nxt
    if infer_box:
        lo, hi = math.inf, 2.2250738585072014e-308
        for p in m.pos:
            lo, hi = min(lo, p[0]), max(hi, p[0])
        ext = hi - lo
        m.box = (ext, ext, ext)
    return m


# ------------------------------------------------------- model -> engine ---
def to_system(m, temperature, rng_state):
    """One-replica engine topology of a model (main.cpp:113-206,372-577):
    positions, types, chains; Go contacts = contacts other than NAT_SS at
    their lengths; disulfide bonds = the NAT_SS contacts; per-bead native
    tether length, angle and dihedral where the model gives a value, and a
    heuristic-term bit where it gives none (angles and dihedrals must join
    consecutive residues of one chain); chirality of every model dihedral
    from the model's positions (after morphing, as the reference sets it
    up)."""
    from .synth import System

    n = m.n_residues
    order = [r for ch in m.chains for r in ch]
    slot = np.empty(n, dtype=np.int64)
    slot[order] = np.arange(n)
    pos = np.array([m.pos[r] for r in order], dtype=np.float64)
    atype = np.array([m.atype[r] for r in order], dtype=np.int8)
    ch_off = np.cumsum([0] + [len(c) for c in m.chains])
    chain_of = np.repeat(np.arange(len(m.chains)), [len(c) for c in m.chains])

    def consecutive(ids):
        s = slot[list(ids)]
        return all(s[k + 1] == s[k] + 1 for k in range(len(s) - 1)) and chain_of[s[0]] == chain_of[s[-1]]

    tether = np.full(n, np.nan)
    for i1, i2, length in m.tethers:
        if not consecutive((i1, i2)):
            raise ValueError(f"tether {i1}-{i2} does not join consecutive residues")
        if length is not None:
            tether[slot[i1]] = length
    theta = np.full(n, np.nan)
    heur = np.zeros(n, dtype=np.uint8)  # angles / dihedrals without a native value
    for *ids, th in m.angles:
        if not consecutive(ids):
            raise ValueError(f"angle {ids} does not join consecutive residues")
        if th is None:
            heur[slot[ids[0]]] |= 1
        else:
            theta[slot[ids[0]]] = th
    phi = np.full(n, np.nan)
    chir = np.full(n, np.nan)
    chir_f = np.full(n, np.nan)
    for *ids, ph in m.dihedrals:
        if not consecutive(ids):
            raise ValueError(f"dihedral {ids} does not join consecutive residues")
        s = slot[ids[0]]
        if ph is not None:
            phi[s] = ph
        else:
            heur[s] |= 2
        r12, r23, r34 = pos[s + 1] - pos[s], pos[s + 2] - pos[s + 1], pos[s + 3] - pos[s + 2]
        f = np.linalg.norm(r23) ** -3.0
        chir[s] = float(r12 @ np.cross(r23, r34)) * f
        chir_f[s] = f
    nat = [(min(slot[i], slot[j]), max(slot[i], slot[j]), d) for i, j, d, t in m.contacts if t != NAT_SS]
    nat.sort()
    nat_i1 = np.array([a for a, _, _ in nat], dtype=np.int32)
    nat_i2 = np.array([b for _, b, _ in nat], dtype=np.int32)
    nat_d = np.array([d for _, _, d in nat], dtype=np.float64)
    ss = sorted((min(slot[i], slot[j]), max(slot[i], slot[j])) for i, j, _, t in m.contacts if t == NAT_SS)

    return System(
        replica_offset=np.array([0, n], dtype=np.int32),
        chain_offset=ch_off.astype(np.int32),
        atype=atype,
        r0=np.ascontiguousarray(pos.T),
        box=np.asarray(m.box, dtype=np.float64).reshape(1, 3),
        temperature=np.array([float(temperature)]),
        rng_state=np.array([int(rng_state) & RandGen.MASK], dtype=np.uint64),
        nat_i1=nat_i1, nat_i2=nat_i2, nat_dist=nat_d,
        tether_len=tether, nat_theta=theta, nat_phi=phi, nat_chir=chir, nat_chir_factor=chir_f,
        ss_i1=np.array([a for a, _ in ss], dtype=np.int32), ss_i2=np.array([b for _, b in ss], dtype=np.int32),
        heur_bonded=heur,
    )


def load_model(tree, base_dir="."):
    """The model of a merged parameter tree's `model` section
    (params.cpp:13-41, main.cpp:113-129) and the generator state after
    morphing: (model, rng_state).  Relative file paths are taken from
    base_dir (the reference takes them from its working directory)."""
    sec = tree["model"]
    gen = RandGen(int(tree["general"]["seed"]))
    if sec.get("seqfile"):
        m = read_seq_file(os.path.join(base_dir, str(sec["seqfile"])))
    elif sec.get("pdb"):
        m = read_pdb(os.path.join(base_dir, str(sec["pdb"])))
    else:
        raise ValueError("model: neither seqfile nor pdb given")
    morph = sec.get("morph into SAW")
    if morph:
        bond = P.quantity(morph["bond distance"]) if morph.get("bond distance") is not None else None
        morph_into_saw(m, gen, bond, P.quantity(morph["residue density"]), P._bool(morph["infer box"]))
    return m, gen.state


def from_input(path_or_dict=None, base_dir=None, replicas=1, **kw):
    """Engine parameters, topology and models of an xmd input file (merged
    over the defaults, as the reference's freeform main does):
    (XmdParams, System, [Model]).  Replica k is the run the reference makes
    with general.seed + k: its own morphed chains and noise stream."""
    from .synth import concat

    tree, base = P.load(path_or_dict)
    base = base if base_dir is None else base_dir
    params = P.to_xmd_params(tree, base, **kw)
    T = P.quantity(tree["langevin"]["temperature"])
    seed = int(tree["general"]["seed"])
    systems, models = [], []
    for k in range(replicas):
        tree["general"]["seed"] = seed + k
        m, state = load_model(tree, base)
        systems.append(to_system(m, T, state))
        models.append(m)
    tree["general"]["seed"] = seed
    system = concat(systems)
    wt = P.walls_and_tips(tree)
    n = [s.n_beads for s in systems]
    off = np.cumsum([0] + n[:-1])
    for key in ("vel_afm", "force_afm"):
        idx = wt.pop(f"{key}_idx")
        if np.any((idx < 0) | (idx >= min(n))):
            raise ValueError(f"{key} tip pulls a residue outside the model")
        setattr(system, f"{key}_idx", np.concatenate([idx + o for o in off]).astype(np.int32))
    for key, val in wt.items():
        setattr(system, key, val if key.endswith("walls") else np.concatenate([val] * len(systems)))
    return params, system, models
