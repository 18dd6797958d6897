"""Seeded synthetic systems for the benchmark configurations C0-C4.

The reference builds its model from seq/PDB files and `morph_into_saw`
(xmd/src/model/model.cpp:101-174); these configurations instead use chains
generated here:
  - native structure: self-avoiding walk, bond 3.8 A, non-bonded beads at
    least 4.5 A apart; residues uniform over the 20 amino acids;
  - native contacts: |i-j| >= 3 and d < 7.5 A within a chain, r_min = d_nat;
  - bonded natives in the conventions of the eval functors: angle
    theta = acos(-r12.r23/|r12||r23|) (xmd/src/forces/angle/native.cpp:29-30),
    dihedral as xmd/src/forces/dihedral/complex_native.cpp:19-28,
    chirality nat_factor = |r23|^-3 and nat_chir = r12.(r23 x r34)*nat_factor
    (xmd/src/freeform/main.cpp:384-391);
  - multi-chain boxes: chains placed at random with every inter-chain pair
    >= 4.5 A under the minimum image.
Seeds: config c uses 0x5EED0000 + c; replica k of a batch uses that + k.
Positions are returned in internal units (5 A).
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np
from scipy.spatial import cKDTree

from . import _abi
from .params import ANGSTROM

BOND = 3.8
MIN_SEP = 4.5
NAT_CUT = 7.5
SEED_BASE = 0x5EED0000
# Largest bond angle of the configurations' native chains (degrees).  A
# self-avoiding walk with uniform bond directions puts ~5 native angles per
# 10k beads within 2 degrees of a straight line, where the native dihedral
# and PID gradients (1/|r12 x r23|) diverge and the reference's force field
# blows up within a few thousand steps; C-alpha traces of proteins have bond
# angles of about 80-150 degrees.
MAX_BOND_ANGLE = 150.0


def saw_chain(n, rng, bond=BOND, min_sep=MIN_SEP, max_tries=200, max_angle=None):
    """Self-avoiding walk of n beads (A), grown bead by bead with
    back-tracking when a bead cannot be placed; with max_angle (degrees)
    every bond angle (at bead k between beads k-1 and k+1) is at most that."""
    pos = np.zeros((n, 3))
    k = 1
    fails = np.zeros(n, dtype=np.int64)
    min2 = min_sep * min_sep
    # bond angle <= max_angle  <=>  b_prev . b_next <= -cos(max_angle) bond^2
    max_dot = None if max_angle is None else -np.cos(np.radians(max_angle)) * bond * bond
    while k < n:
        placed = False
        for _ in range(max_tries):
            d = rng.normal(size=3)
            d *= bond / np.linalg.norm(d)
            cand = pos[k - 1] + d
            if max_dot is not None and k >= 2 and np.dot(pos[k - 1] - pos[k - 2], d) > max_dot:
                continue
            if k >= 2:
                diff = pos[:k - 1] - cand
                if np.min(np.einsum("ij,ij->i", diff, diff)) < min2:
                    continue
            pos[k] = cand
            placed = True
            break
        if placed:
            k += 1
        else:
            fails[k] += 1
            k = max(1, k - 1 - int(fails[k]))
    return pos


def native_contacts(pos, cut=NAT_CUT, min_sep=3):
    tree = cKDTree(pos)
    pairs = tree.query_pairs(cut, output_type="ndarray")
    if len(pairs) == 0:
        return np.zeros((0, 2), dtype=np.int64), np.zeros(0)
    pairs = np.sort(pairs, axis=1)
    pairs = pairs[(pairs[:, 1] - pairs[:, 0]) >= min_sep]
    order = np.lexsort((pairs[:, 1], pairs[:, 0]))
    pairs = pairs[order]
    d = np.linalg.norm(pos[pairs[:, 1]] - pos[pairs[:, 0]], axis=1)
    return pairs, d


def bonded_natives(pos):
    """Per-bead native angle, dihedral and chirality of one chain (A)."""
    n = len(pos)
    theta = np.full(n, np.nan)
    phi = np.full(n, np.nan)
    chir = np.full(n, np.nan)
    chir_f = np.full(n, np.nan)
    if n >= 3:
        r12 = pos[1:-1] - pos[:-2]
        r23 = pos[2:] - pos[1:-1]
        c = -np.einsum("ij,ij->i", r12, r23) / (np.linalg.norm(r12, axis=1) * np.linalg.norm(r23, axis=1))
        theta[:n - 2] = np.arccos(np.clip(c, -1, 1))
    if n >= 4:
        r12 = pos[1:-2] - pos[:-3]
        r23 = pos[2:-1] - pos[1:-2]
        r34 = pos[3:] - pos[2:-1]
        x12_23 = np.cross(r12, r23)
        x23_34 = np.cross(r23, r34)
        u1 = x12_23 / np.linalg.norm(x12_23, axis=1)[:, None]
        u2 = x23_34 / np.linalg.norm(x23_34, axis=1)[:, None]
        p = np.arccos(np.clip(np.einsum("ij,ij->i", u1, u2), -1, 1))
        p = np.where(np.einsum("ij,ij->i", x12_23, r34) < 0, -p, p)
        phi[:n - 3] = p
        f = np.linalg.norm(r23, axis=1) ** -3.0
        chir[:n - 3] = np.einsum("ij,ij->i", r12, x23_34) * f
        chir_f[:n - 3] = f
    return theta, phi, chir, chir_f


@dataclass
class System:
    """Concatenated replicas in the layout of xmd_topology (internal units)."""
    replica_offset: np.ndarray
    chain_offset: np.ndarray
    atype: np.ndarray
    r0: np.ndarray                 # (3, N)
    box: np.ndarray                # (R, 3)
    temperature: np.ndarray        # (R,)
    rng_state: np.ndarray          # (R,) uint64
    nat_i1: np.ndarray
    nat_i2: np.ndarray
    nat_dist: np.ndarray
    tether_len: np.ndarray
    nat_theta: np.ndarray
    nat_phi: np.ndarray
    nat_chir: np.ndarray
    nat_chir_factor: np.ndarray
    ss_i1: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    ss_i2: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    # wall planes (k, 6): origin, unit normal; the same for every replica
    solid_walls: np.ndarray = field(default_factory=lambda: np.zeros((0, 6)))
    lj_walls: np.ndarray = field(default_factory=lambda: np.zeros((0, 6)))
    # AFM tips: global bead index, origin (NaN = initial position) and
    # velocity, or constant force
    vel_afm_idx: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    vel_afm_orig: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    vel_afm_v: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    force_afm_idx: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    force_afm_F: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    # per bead or None: bit 0 angle i..i+2, bit 1 dihedral i..i+3 in the
    # model without a native value (the heuristic terms' entries)
    heur_bonded: np.ndarray = None
    _keep: list = field(default_factory=list, repr=False)

    @property
    def n_beads(self):
        return int(self.replica_offset[-1])

    @property
    def n_replicas(self):
        return len(self.replica_offset) - 1

    def replica_slice(self, k):
        return slice(int(self.replica_offset[k]), int(self.replica_offset[k + 1]))

    def topology(self):
        """XmdTopology pointing into this system's (contiguous) arrays."""
        arrs = dict(
            replica_offset=np.ascontiguousarray(self.replica_offset, dtype=np.int32),
            chain_offset=np.ascontiguousarray(self.chain_offset, dtype=np.int32),
            atype=np.ascontiguousarray(self.atype, dtype=np.int8),
            r0=np.ascontiguousarray(self.r0, dtype=np.float64).reshape(-1),
            box=np.ascontiguousarray(self.box, dtype=np.float64).reshape(-1),
            temperature=np.ascontiguousarray(self.temperature, dtype=np.float64),
            rng_state=np.ascontiguousarray(self.rng_state, dtype=np.uint64),
            nat_i1=np.ascontiguousarray(self.nat_i1, dtype=np.int32),
            nat_i2=np.ascontiguousarray(self.nat_i2, dtype=np.int32),
            nat_dist=np.ascontiguousarray(self.nat_dist, dtype=np.float64),
            tether_len=np.ascontiguousarray(self.tether_len, dtype=np.float64),
            nat_theta=np.ascontiguousarray(self.nat_theta, dtype=np.float64),
            nat_phi=np.ascontiguousarray(self.nat_phi, dtype=np.float64),
            nat_chir=np.ascontiguousarray(self.nat_chir, dtype=np.float64),
            nat_chir_factor=np.ascontiguousarray(self.nat_chir_factor, dtype=np.float64),
            ss_i1=np.ascontiguousarray(self.ss_i1, dtype=np.int32),
            ss_i2=np.ascontiguousarray(self.ss_i2, dtype=np.int32),
            solid_walls=np.ascontiguousarray(self.solid_walls, dtype=np.float64).reshape(-1),
            lj_walls=np.ascontiguousarray(self.lj_walls, dtype=np.float64).reshape(-1),
            vel_afm_idx=np.ascontiguousarray(self.vel_afm_idx, dtype=np.int32),
            vel_afm_orig=np.ascontiguousarray(self.vel_afm_orig, dtype=np.float64).reshape(-1),
            vel_afm_v=np.ascontiguousarray(self.vel_afm_v, dtype=np.float64).reshape(-1),
            force_afm_idx=np.ascontiguousarray(self.force_afm_idx, dtype=np.int32),
            force_afm_F=np.ascontiguousarray(self.force_afm_F, dtype=np.float64).reshape(-1),
            heur_bonded=None if self.heur_bonded is None else np.ascontiguousarray(self.heur_bonded, dtype=np.uint8),
        )
        t = _abi.XmdTopology()
        t.n_replicas = self.n_replicas
        t.n_beads = self.n_beads
        t.n_chains = len(self.chain_offset) - 1
        t.n_native = len(self.nat_i1)
        t.n_ssbonds = len(self.ss_i1)
        t.n_solid_walls = len(self.solid_walls)
        t.n_lj_walls = len(self.lj_walls)
        t.n_vel_afm = len(self.vel_afm_idx)
        t.n_force_afm = len(self.force_afm_idx)
        types = dict(replica_offset=C.c_int32, chain_offset=C.c_int32, atype=C.c_int8,
                     rng_state=C.c_uint64, nat_i1=C.c_int32, nat_i2=C.c_int32, ss_i1=C.c_int32, ss_i2=C.c_int32,
                     vel_afm_idx=C.c_int32, force_afm_idx=C.c_int32, heur_bonded=C.c_uint8)
        for name, a in arrs.items():
            setattr(t, name, _abi.ptr(a, types.get(name, C.c_double)))
        t._arrays = arrs  # keep alive with the struct
        return t


def _wrap(x, L):
    return x - L * np.floor(x / L) if L > 0 else x


def place_chains(chains, L, rng, min_sep=MIN_SEP, max_tries=2000):
    """Translate chains (A) to random positions in a periodic cube of side L
    with all inter-chain pairs at least min_sep apart."""
    placed = []
    tree = None
    allpos = np.zeros((0, 3))
    for ch in chains:
        c0 = ch - ch.mean(axis=0)
        for _ in range(max_tries):
            shift = rng.uniform(0, L, size=3)
            cand = c0 + shift
            if tree is not None:
                d, _ = tree.query(_wrap(cand, L), k=1)
                if np.min(d) < min_sep:
                    continue
            break
        else:
            raise RuntimeError("could not place chain without overlap")
        placed.append(cand)
        allpos = np.vstack([allpos, cand])
        tree = cKDTree(_wrap(allpos, L), boxsize=L)
    return placed


def make_system(replicas):
    """replicas: list of dicts with keys chains (list of (pos_A, atype)),
    box_A (3,), temperature, seed."""
    rep_off, ch_off = [0], [0]
    atype, pos, nat_i1, nat_i2, nat_d = [], [], [], [], []
    theta, phi, chir, chir_f = [], [], [], []
    box, temp, seeds = [], [], []
    n = 0
    for rep in replicas:
        for p, at in rep["chains"]:
            m = len(p)
            atype.append(np.asarray(at, dtype=np.int8))
            pos.append(p)
            pairs, d = native_contacts(p)
            nat_i1.append(pairs[:, 0] + n)
            nat_i2.append(pairs[:, 1] + n)
            nat_d.append(d)
            th, ph, cc, cf = bonded_natives(p)
            theta.append(th)
            phi.append(ph)
            chir.append(cc)
            chir_f.append(cf)
            n += m
            ch_off.append(n)
        rep_off.append(n)
        box.append(np.asarray(rep["box_A"], dtype=np.float64))
        temp.append(rep["temperature"])
        seeds.append(rep["seed"])
    pos = np.vstack(pos)
    return System(
        replica_offset=np.array(rep_off, dtype=np.int32),
        chain_offset=np.array(ch_off, dtype=np.int32),
        atype=np.concatenate(atype),
        r0=np.ascontiguousarray(pos.T * ANGSTROM),
        box=np.array(box) * ANGSTROM,
        temperature=np.array(temp, dtype=np.float64),
        rng_state=np.array(seeds, dtype=np.uint64),
        nat_i1=np.concatenate(nat_i1).astype(np.int32),
        nat_i2=np.concatenate(nat_i2).astype(np.int32),
        nat_dist=np.concatenate(nat_d) * ANGSTROM,
        tether_len=np.full(n, np.nan),
        nat_theta=np.concatenate(theta),
        nat_phi=np.concatenate(phi),
        nat_chir=np.concatenate(chir),
        nat_chir_factor=np.concatenate(chir_f) / ANGSTROM ** 3,
    )


def random_chain(n, rng, max_angle=None):
    return saw_chain(n, rng, max_angle=max_angle), rng.integers(0, 20, size=n)


# Term sets of the configurations (names of the YAML sections' enabled flags).
C0_TERMS = dict(go=True, pauli=True, tether=True, chir=True, nat_ang=False, cnd=False, snd=False,
                const_es=False, rel_es=False, qa=False, pid=False, nat_ss=False)
C1_TERMS = dict(C0_TERMS, const_es=True, nat_ang=True, cnd=True)
FULL_TERMS = dict(go=True, pauli=True, tether=True, chir=True, nat_ang=True, cnd=True, snd=False,
                  const_es=True, rel_es=False, qa=True, pid=True, nat_ss=False)
C3_TERMS = dict(C0_TERMS, const_es=True)

_FLAG = dict(go="go_enabled", pauli="pauli_enabled", tether="tether_enabled", chir="chir_enabled",
             nat_ang="nat_ang_enabled", cnd="cnd_enabled", snd="snd_enabled", const_es="const_es_enabled",
             rel_es="rel_es_enabled", qa="qa_enabled", pid="pid_enabled", nat_ss="nat_ss_enabled",
             solid_wall="solid_walls_enabled", lj_wall="lj_walls_enabled", vel_afm="vel_afm_enabled",
             force_afm="force_afm_enabled", heur_ang="heur_ang_enabled", heur_dih="heur_dih_enabled")


def apply_terms(p, terms):
    for k, v in terms.items():
        setattr(p, _FLAG[k], int(bool(v)))
    return p


def config_params(cid, precision=_abi.PREC_FP64, rng_exact=True):
    """XmdParams of configuration cid (YAML defaults + the config's terms)."""
    from .params import from_yaml
    p, _ = from_yaml(None, precision=precision, rng_exact=rng_exact)
    p.dt = 0.005
    p.gamma_factor = 2.0
    terms = {0: C0_TERMS, 1: C1_TERMS, 2: FULL_TERMS, 3: C3_TERMS, 4: FULL_TERMS}[cid]
    apply_terms(p, terms)
    if cid == 0:  # "Verlet cutoff 18 A + 2 A pad"
        p.max_cutoff_override = 18.0 * ANGSTROM
        p.pad_factor = 2.0 / 18.0
    return p


def _pick_seeds(start, count, n_beads_of, horizon):
    """`count` seeds from `start` upward, skipping any whose rand_gen stream
    feeds a zero uniform draw to log() within `horizon` steps: that draw
    makes the reference's Langevin noise infinite (rand_gen.h:45-46)."""
    if horizon <= 0:
        return [start + k for k in range(count)]
    from .engine import first_zero_log_draw
    seeds, cand = [], start
    while len(seeds) < count:
        if first_zero_log_draw(cand, n_beads_of(len(seeds)), horizon) is None:
            seeds.append(cand)
        cand += 1
    return seeds


def config_system(cid, n_batch=1, n_replicas=None, chains=None, chain_len=None, seed=None, horizon_steps=0,
                  max_angle=MAX_BOND_ANGLE):
    """Synthetic system of configuration cid.

    C0: 1 x 60, open box.  C1: 1 x 1000, open box.
    C2: 40 x 250 in a periodic cube at 1e-4 residues/A^3 (L = 464.2 A).
    C3: 512 single chains, lengths U{30..400}, T ~ U[0.30, 0.50].
    C4: 400 x 250 at 5e-4 residues/A^3 (L = 584.8 A).
    n_batch > 1 repeats C2/C4 as independent boxes with temperatures
    0.30 + 0.2 k/(n_batch-1) and seeds base + k.  With horizon_steps > 0,
    seeds whose noise stream hits a zero log() draw within that many steps
    are skipped (the next seed is taken).  Native chains have bond angles of
    at most max_angle degrees (None: any).
    """
    base = SEED_BASE + cid if seed is None else seed
    reps = []
    if cid in (0, 1):
        n = {0: 60, 1: 1000}[cid] if chain_len is None else chain_len
        for k, sd in enumerate(_pick_seeds(base, n_batch, lambda _: n, horizon_steps)):
            rng = np.random.default_rng(sd)
            reps.append(dict(chains=[random_chain(n, rng, max_angle)], box_A=(0, 0, 0), temperature=0.35, seed=sd))
    elif cid in (2, 4):
        nch = {2: 40, 4: 400}[cid] if chains is None else chains
        m = 250 if chain_len is None else chain_len
        dens = {2: 1e-4, 4: 5e-4}[cid]
        L = (nch * m / dens) ** (1.0 / 3.0)
        for k, sd in enumerate(_pick_seeds(base, n_batch, lambda _: nch * m, horizon_steps)):
            rng = np.random.default_rng(sd)
            chs = [random_chain(m, rng, max_angle) for _ in range(nch)]
            placed = place_chains([c[0] for c in chs], L, rng)
            T = 0.35 if n_batch == 1 else 0.30 + 0.2 * k / (n_batch - 1)
            reps.append(dict(chains=[(p, c[1]) for p, c in zip(placed, chs)], box_A=(L, L, L),
                             temperature=T, seed=sd))
    elif cid == 3:
        nrep = 512 if n_replicas is None else n_replicas
        rng0 = np.random.default_rng(base)
        lens = rng0.integers(30, 401, size=nrep)
        temps = rng0.uniform(0.30, 0.50, size=nrep)
        for k, sd in enumerate(_pick_seeds(base + 1, nrep, lambda i: int(lens[i]), horizon_steps)):
            rng = np.random.default_rng(sd)
            reps.append(dict(chains=[random_chain(int(lens[k]), rng, max_angle)], box_A=(0, 0, 0),
                             temperature=float(temps[k]), seed=sd))
    else:
        raise ValueError(cid)
    return make_system(reps)


def replicate(system, temperatures, seeds):
    """len(temperatures) copies of a one-replica `system` (same chains and
    initial positions) as replicas of one batch, each with its own
    temperature and rand_gen seed."""
    assert system.n_replicas == 1
    n, k = system.n_beads, len(temperatures)
    tile = lambda a: np.concatenate([a] * k)
    shifted = lambda a: np.concatenate([a + q * n for q in range(k)])
    nch = len(system.chain_offset) - 1
    return System(
        replica_offset=np.arange(k + 1, dtype=np.int32) * n,
        chain_offset=np.concatenate([system.chain_offset[:nch] + q * n for q in range(k)] + [[k * n]]).astype(np.int32),
        atype=tile(system.atype),
        r0=np.concatenate([system.r0] * k, axis=1),
        box=np.repeat(system.box.reshape(1, 3), k, axis=0),
        temperature=np.asarray(temperatures, dtype=np.float64),
        rng_state=np.asarray(seeds, dtype=np.uint64),
        nat_i1=shifted(system.nat_i1), nat_i2=shifted(system.nat_i2), nat_dist=tile(system.nat_dist),
        tether_len=tile(system.tether_len), nat_theta=tile(system.nat_theta), nat_phi=tile(system.nat_phi),
        nat_chir=tile(system.nat_chir), nat_chir_factor=tile(system.nat_chir_factor),
        ss_i1=shifted(system.ss_i1).astype(np.int32), ss_i2=shifted(system.ss_i2).astype(np.int32),
        solid_walls=system.solid_walls, lj_walls=system.lj_walls,
        vel_afm_idx=shifted(system.vel_afm_idx).astype(np.int32),
        vel_afm_orig=np.concatenate([system.vel_afm_orig.reshape(-1, 3)] * k),
        vel_afm_v=np.concatenate([system.vel_afm_v.reshape(-1, 3)] * k),
        force_afm_idx=shifted(system.force_afm_idx).astype(np.int32),
        force_afm_F=np.concatenate([system.force_afm_F.reshape(-1, 3)] * k),
        heur_bonded=None if system.heur_bonded is None else tile(system.heur_bonded))


def concat(systems):
    """Systems (of any replica counts) as the replicas of one batch, in order."""
    off = np.cumsum([0] + [s.n_beads for s in systems])
    cat = lambda name: np.concatenate([getattr(s, name) for s in systems])
    shift = lambda name: np.concatenate([getattr(s, name) + o for s, o in zip(systems, off)])
    return System(
        replica_offset=np.concatenate([s.replica_offset[:-1] + o for s, o in zip(systems, off)] + [[off[-1]]])
        .astype(np.int32),
        chain_offset=np.concatenate([s.chain_offset[:-1] + o for s, o in zip(systems, off)] + [[off[-1]]])
        .astype(np.int32),
        atype=cat("atype"), r0=np.concatenate([s.r0 for s in systems], axis=1),
        box=np.concatenate([s.box.reshape(-1, 3) for s in systems]), temperature=cat("temperature"),
        rng_state=cat("rng_state").astype(np.uint64),
        nat_i1=shift("nat_i1").astype(np.int32), nat_i2=shift("nat_i2").astype(np.int32), nat_dist=cat("nat_dist"),
        tether_len=cat("tether_len"), nat_theta=cat("nat_theta"), nat_phi=cat("nat_phi"),
        nat_chir=cat("nat_chir"), nat_chir_factor=cat("nat_chir_factor"),
        ss_i1=shift("ss_i1").astype(np.int32), ss_i2=shift("ss_i2").astype(np.int32),
        solid_walls=systems[0].solid_walls, lj_walls=systems[0].lj_walls,
        vel_afm_idx=shift("vel_afm_idx").astype(np.int32),
        vel_afm_orig=np.concatenate([s.vel_afm_orig.reshape(-1, 3) for s in systems]),
        vel_afm_v=np.concatenate([s.vel_afm_v.reshape(-1, 3) for s in systems]),
        force_afm_idx=shift("force_afm_idx").astype(np.int32),
        force_afm_F=np.concatenate([s.force_afm_F.reshape(-1, 3) for s in systems]),
        heur_bonded=None if all(s.heur_bonded is None for s in systems) else np.concatenate(
            [np.zeros(s.n_beads, np.uint8) if s.heur_bonded is None else s.heur_bonded for s in systems]))


def subset(system, replicas):
    """The replicas `replicas` (in that order) of a batch as a batch of
    their own."""
    parts = []
    for k in replicas:
        sl = system.replica_slice(k)
        b0, n = sl.start, sl.stop - sl.start
        m = (system.nat_i1 >= b0) & (system.nat_i1 < sl.stop)
        ms = (system.ss_i1 >= b0) & (system.ss_i1 < sl.stop)
        co = system.chain_offset
        ch = co[(co >= b0) & (co <= sl.stop)] - b0
        hb = None if system.heur_bonded is None else system.heur_bonded[sl]
        parts.append(System(
            replica_offset=np.array([0, n], dtype=np.int32), chain_offset=ch.astype(np.int32),
            atype=system.atype[sl], r0=system.r0[:, sl], box=system.box.reshape(-1, 3)[k:k + 1],
            temperature=system.temperature[k:k + 1], rng_state=system.rng_state[k:k + 1],
            nat_i1=(system.nat_i1[m] - b0).astype(np.int32), nat_i2=(system.nat_i2[m] - b0).astype(np.int32),
            nat_dist=system.nat_dist[m], tether_len=system.tether_len[sl], nat_theta=system.nat_theta[sl],
            nat_phi=system.nat_phi[sl], nat_chir=system.nat_chir[sl], nat_chir_factor=system.nat_chir_factor[sl],
            ss_i1=(system.ss_i1[ms] - b0).astype(np.int32), ss_i2=(system.ss_i2[ms] - b0).astype(np.int32),
            solid_walls=system.solid_walls, lj_walls=system.lj_walls, heur_bonded=hb))
    return concat(parts)


def perturbed_positions(system, seed, sigma_A=0.3):
    """Native positions plus isotropic Gaussian displacement (C1's seeded
    conformations)."""
    rng = np.random.default_rng(seed)
    return system.r0 + rng.normal(scale=sigma_A * ANGSTROM, size=system.r0.shape)
