"""Model loaders (sequence files with contact maps, PDB files, morph into a
random chain) and the topology built from a model.

Inputs are the reference's own example models (xmd/data/examples, copied to
tests/golden/models); expected values are read off those files by hand
(literal numbers below) or follow from closed forms: the generator's draw
count, bond lengths, tilt bounds.  The oracle's independent restatement of
the random-chain layout (Rodrigues matrices against the loader's
quaternions) pins the positions.
"""
import math
import os

import numpy as np
import pytest

from xmd_amd import model as M, params as P
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
MODELS = os.path.join(HERE, "golden", "models")


@pytest.fixture(scope="module")
def glut():
    return M.read_seq_file(os.path.join(MODELS, "glut", "glut.yml"))


def test_seq_file_chains_and_maps(glut):
    # 14 chains: xhm, yhm, agli, lmw, ygli parts as listed in glut.yml
    assert [len(c) for c in glut.chains] == [827, 827, 627, 287, 287, 287, 284, 284, 284, 284, 272, 272, 272, 272]
    # contacts: per chain, its maps' contacts (xhmn 205 + xhmc 93 for xhm ...)
    assert len(glut.contacts) == 6209
    # first contact of xhmn.map.yml, then xhmc's first one shifted by 785
    assert glut.contacts[0][:3] == (0, 41, 1.3895906584853075)
    assert (785, 788, 1.27153525710619, M.UNKNOWN) in glut.contacts
    # xhmc's first angle "0,1,2,1.4245863345954886" shifted by 785
    assert (785, 786, 787, 1.4245863345954886) in glut.angles
    # every consecutive triple and quad once: map values or none
    n = glut.n_residues
    assert len(glut.angles) == sum(len(c) - 2 for c in glut.chains)
    assert len(glut.dihedrals) == sum(len(c) - 3 for c in glut.chains)
    assert len({a[:3] for a in glut.angles}) == len(glut.angles)
    assert len(glut.tethers) == n - len(glut.chains)
    assert all(t[2] is None for t in glut.tethers)


def test_seq_file_types():
    m = M.read_seq_file(os.path.join(MODELS, "9aac", "9aac.yml"))
    assert len(m.chains) == 1 and len(m.chains[0]) == 140
    # "MDVFMKGLSK..." in aa_code order ALA ARG ASN ASP CYS GLU GLN GLY HIS ILE LEU LYS MET PHE PRO ...
    assert m.atype[:10] == [12, 3, 19, 13, 12, 11, 7, 10, 15, 11]
    assert m.contacts == [] and all(a[3] is None for a in m.angles)


def test_pdb_model():
    m = M.read_pdb(os.path.join(MODELS, "1ubq", "1ubq.pdb"))
    assert [len(c) for c in m.chains] == [76]
    A = P.ANGSTROM
    # CRYST1   50.840   42.770   28.950
    assert m.box == pytest.approx((50.840 * A, 42.770 * A, 28.950 * A), rel=1e-15)
    # ATOM 2 CA MET A 1 26.266 25.413 2.842 ; ATOM 599 CA GLY A 76 40.373 39.813 33.944
    assert m.pos[0] == pytest.approx((26.266 * A, 25.413 * A, 2.842 * A), rel=1e-15)
    assert m.pos[75] == pytest.approx((40.373 * A, 39.813 * A, 33.944 * A), rel=1e-15)
    assert m.atype[0] == P.AA_INDEX["MET"] and m.atype[75] == P.AA_INDEX["GLY"]
    assert m.contacts == []  # no SSBOND or LINK records
    p = np.array(m.pos)
    # tethers at the CA distance (~3.8 A)
    for i1, i2, length in m.tethers:
        assert length == pytest.approx(np.linalg.norm(p[i2] - p[i1]), rel=1e-15)
        assert 3.7 * A < length < 3.9 * A
    # angle between the bond vectors (the supplement of the CA-CA-CA angle)
    i1, i2, i3, th = m.angles[0]
    u, v = p[i2] - p[i1], p[i3] - p[i2]
    bond_angle = math.acos(float(-(u @ v)) / (np.linalg.norm(u) * np.linalg.norm(v)))
    assert th == pytest.approx(math.pi - bond_angle, abs=1e-12)
    # dihedral sign from (r12 x r23) . r34
    i1, i2, i3, i4, ph = m.dihedrals[0]
    r12, r23, r34 = p[i2] - p[i1], p[i3] - p[i2], p[i4] - p[i3]
    x1, x2 = np.cross(r12, r23), np.cross(r23, r34)
    assert abs(ph) == pytest.approx(math.acos(float(x1 @ x2) / (np.linalg.norm(x1) * np.linalg.norm(x2))), abs=1e-12)
    assert math.copysign(1.0, ph) == math.copysign(1.0, float(x1 @ r34))


PDB_SNIPPET = """\
CRYST1   20.000   30.000   40.000  90.00  90.00  90.00 P 1           1
SSBOND   1 CYS A    2    CYS B    1                          1555   1555  2.03
LINK         NZ  LYS A   3                 O   GLY B   2     1555   1555  2.50
ATOM      1  CA  ALA A   1       0.000   0.000   0.000  1.00  0.00           C
ATOM      2  CA  CYS A   2       3.800   0.000   0.000  1.00  0.00           C
ATOM      3  N   LYS A   3       7.000   0.500   0.000  1.00  0.00           N
ATOM      4  CA  LYS A   3       7.600   0.000   0.000  1.00  0.00           C
ATOM      5  NZ  LYS A   3       9.000   2.000   0.000  1.00  0.00           N
HETATM    6  O   HOH A 101       1.000   1.000   1.000  1.00  0.00           O
TER       7      LYS A   3
ATOM      8  CA  CYS B   1       0.000   5.000   0.000  1.00  0.00           C
ATOM      9  CA  GLY B   2       3.800   5.000   0.000  1.00  0.00           C
ATOM     10  O   GLY B   2       4.500   6.000   0.000  1.00  0.00           O
END{pad}
ATOM     11  CA  ALA C   1       0.000   0.000   9.000  1.00  0.00           C
"""


def test_pdb_records_contacts_and_end():
    # columns 1-6 must read "END   ": a bare "END" does not end the file
    assert [len(c) for c in M.read_pdb(PDB_SNIPPET.format(pad="")).chains] == [3, 2, 1]
    m = M.read_pdb(PDB_SNIPPET.format(pad="   "))
    A = P.ANGSTROM
    assert [len(c) for c in m.chains] == [3, 2]  # HETATM ignored, nothing after END
    assert m.box == pytest.approx((20 * A, 30 * A, 40 * A))
    # SSBOND A2-B1 -> NAT_SS; LINK A3 NZ (side) - B2 O (backbone) -> SIDE_BACK
    ss = [c for c in m.contacts if c[3] == M.NAT_SS]
    link = [c for c in m.contacts if c[3] != M.NAT_SS]
    assert [(c[0], c[1]) for c in ss] == [(1, 3)]
    assert ss[0][2] == pytest.approx(math.hypot(3.8, 5.0) * A)
    assert [(c[0], c[1], c[3]) for c in link] == [(2, 4, M.SIDE_BACK)]
    assert link[0][2] == pytest.approx(math.hypot(3.8, 5.0) * A)
    # the Go contacts of the topology leave out the SS bond
    s = M.to_system(m, 0.35, 1)
    assert list(zip(s.nat_i1, s.nat_i2)) == [(2, 4)]
    assert list(zip(s.ss_i1, s.ss_i2)) == [(1, 3)]


def test_rand_gen_matches_oracle():
    g, st = M.RandGen(0x5EED0001), 0x5EED0001
    for _ in range(1000):
        v, st = O.rng_next(st)
        assert g.next() == v
    assert g.state == st
    # uniform from the top 32 bits
    g2 = M.RandGen(7)
    raw = M.RandGen(7).next()
    assert g2.uniform() == (raw >> 32) / 2.0 ** 32


@pytest.mark.parametrize("bond", [0.76, None])
def test_morph_into_saw(glut, bond):
    m = M.read_seq_file(os.path.join(MODELS, "glut", "glut.yml")) if bond else M.read_pdb(
        os.path.join(MODELS, "1ubq", "1ubq.pdb"))
    orig = np.array(m.pos)
    lens = [len(c) for c in m.chains]
    g = M.RandGen(42)
    M.morph_into_saw(m, g, bond, 0.125, True)
    p = np.array(m.pos)
    # draws: 3 uniforms + sphere (4) per chain, 2 per bond
    ref = M.RandGen(42)
    for _ in range(sum(7 + 2 * (L - 1) for L in lens)):
        ref.next()
    assert g.state == ref.state
    # the oracle's independent layout (Rodrigues matrices)
    po, box, st = O.morph_into_saw(42, lens, orig, bond, 0.125, True)
    assert st == g.state
    assert np.abs(p - po).max() < 1e-11 * max(1.0, np.abs(po).max())
    assert m.box == pytest.approx(tuple(box), rel=1e-15)
    a = (sum(lens) / 0.125) ** (1.0 / 3.0)
    start = 0
    for L in lens:
        c = p[start:start + L]
        assert np.all(np.abs(c[0]) <= a / 2)
        d = np.diff(c, axis=0)
        bl = np.linalg.norm(d, axis=1)
        want = np.full(L - 1, bond) if bond else np.linalg.norm(np.diff(orig[start:start + L], axis=0), axis=1)
        assert np.allclose(bl, want, rtol=1e-12)
        u = d / bl[:, None]
        cosang = np.einsum("ij,ij->i", u[1:], u[:-1])
        assert np.all(cosang >= math.cos(math.pi / 12) - 1e-12)  # tilt at most 15 degrees
        start += L
    # box: every extent from the x coordinates
    assert m.box[0] == m.box[1] == m.box[2] == pytest.approx(p[:, 0].max() - p[:, 0].min())


def test_input_file_topology():
    """Defaults (glut sequence file, morph with seed 42) through the input
    front end: topology arrays and the generator state for the noise."""
    params, system, models = M.from_input({"model": {"seqfile": "glut/glut.yml"}}, base_dir=MODELS, replicas=2)
    m = models[0]
    n = m.n_residues
    assert system.n_replicas == 2 and system.n_beads == 2 * n
    assert len(system.chain_offset) - 1 == 2 * len(m.chains)
    # Go contacts: all 6209 map contacts per replica, i1 < i2, raw lengths
    assert len(system.nat_i1) == 2 * 6209
    assert np.all(system.nat_i1 < system.nat_i2)
    assert np.nanmax(system.nat_dist) == pytest.approx(1.4252299201301868)
    # native angles / dihedrals where a map gives one, chirality everywhere a quad exists
    assert np.sum(~np.isnan(system.nat_theta[:n])) == sum(a[3] is not None for a in m.angles)
    assert np.sum(~np.isnan(system.nat_phi[:n])) == sum(d[4] is not None for d in m.dihedrals)
    assert np.sum(~np.isnan(system.nat_chir[:n])) == n - 3 * len(m.chains)
    assert np.all(np.isnan(system.tether_len))
    # replica k = seed 42 + k, noise starts after the morph draws
    for k in range(2):
        g = M.RandGen(42 + k)
        for _ in range(sum(7 + 2 * (len(c) - 1) for c in m.chains)):
            g.next()
        assert int(system.rng_state[k]) == g.state
    assert not np.allclose(system.r0[:, :n], system.r0[:, n:])
    assert system.temperature[0] == pytest.approx(0.38)
    # the oracle takes the topology and steps it
    from oracle.oracle import Oracle
    ora = Oracle(params, system)
    ora.step(3)
    F, V = ora.eval_forces()
    assert np.all(np.isfinite(F))
