"""Heuristic angle and dihedral terms (xmd/src/forces/angle/heurestic.cpp,
xmd/src/forces/dihedral/heurestic.cpp; tables params.cpp:69-124), pinned on
the oracle: the tables as the defaults hold them, closed forms on hand-built
geometries (numpy's polyval and math.sin/cos, not the oracle's Horner
sweep), every pair type, finite-difference gradients (which also pin D30's
exact dV/dphi), Newton's third law, and the native value taking precedence.
"""
import dataclasses
import math

import numpy as np
import pytest

from xmd_amd import _abi, params as P, synth
from oracle.oracle import Oracle

A = synth.ANGSTROM
GLY, PRO, ALA = P.AA_INDEX["GLY"], P.AA_INDEX["PRO"], P.AA_INDEX["ALA"]
KA, KD = _abi.TERM["heur_ang"], _abi.TERM["heur_dih"]


def heur_params(angles=True, dihedrals=True):
    p, _ = P.from_yaml(None)
    for name in _abi.TERMS:
        flag = synth._FLAG.get({"chiral": "chir"}.get(name, name))
        if flag:
            setattr(p, flag, 0)
    p.heur_ang_enabled, p.heur_dih_enabled = int(angles), int(dihedrals)
    p.langevin_enabled = 0
    return p


def chain_system(pos_A, atype, heur=3):
    """One open chain with no native angles or dihedrals and the heuristic
    bits set on every interior triple (bit 0) and quad (bit 1)."""
    pos_A = np.asarray(pos_A, dtype=np.float64)
    n = len(pos_A)
    base = synth.make_system([dict(chains=[(pos_A, list(atype))], box_A=(0, 0, 0), temperature=0.35, seed=3)])
    bits = np.zeros(n, dtype=np.uint8)
    if heur & 1:
        bits[:n - 2] |= 1
    if heur & 2:
        bits[:n - 3] |= 2
    return dataclasses.replace(base, nat_theta=np.full(n, np.nan), nat_phi=np.full(n, np.nan), heur_bonded=bits)


def test_tables_read_by_position():
    tree, _ = P.load(None)
    ang, dih = P.heuristic_tables(tree)
    assert ang[3 * 0 + 0] == pytest.approx([20872.75597, -63260.52963, 79318.301, -52680.17088, 19554.21897,
                                             -3847.670279, 313.6716916], rel=0, abs=0)
    assert ang[3 * 2 + 1, 6] == 325.3085829          # X,P row, 7th number
    assert dih[0, 0] == 0.133672207 and dih[0, 5] == -0.060606303  # G,G (trailing comma)
    assert dih[3 * 2 + 1, 0] == 0.198889776          # X,P listed before X,G
    assert dih[3 * 2 + 0, 5] == -0.790909154
    # angles near 110 deg cost a few eps with all 7 coefficients
    for k in range(9):
        assert 0 < np.polynomial.polynomial.polyval(1.9, ang[k]) < 5
    with pytest.raises(ValueError):
        P.heuristic_tables({"heurestic angles": {"coefficients": "type1,type2,a0\nG,G,1,2,3\n"}})


@pytest.mark.parametrize("a1,a2", [(GLY, GLY), (GLY, PRO), (PRO, ALA), (ALA, GLY), (ALA, ALA)])
def test_angle_closed_form_every_type(a1, a2):
    p = heur_params(dihedrals=False)
    tree, _ = P.load(None)
    ang, _ = P.heuristic_tables(tree)
    cls = {GLY: 0, PRO: 1}
    L = 3.8
    for theta in (1.7, 2.0, 2.3):
        pos = [[L, 0, 0], [0, 0, 0], [L * math.cos(theta), L * math.sin(theta), 0]]
        ora = Oracle(p, chain_system(pos, [a1, a2, ALA]))
        _, V = ora.eval_forces(1 << KA)
        want = np.polynomial.polynomial.polyval(theta, ang[3 * cls.get(a1, 2) + cls.get(a2, 2)])
        assert V[0, KA] == pytest.approx(want, rel=1e-9, abs=1e-9)


@pytest.mark.parametrize("a2,a3", [(PRO, GLY), (GLY, PRO), (ALA, PRO), (ALA, ALA)])
def test_dihedral_closed_form_every_type(a2, a3):
    p = heur_params(angles=False)
    tree, _ = P.load(None)
    _, dih = P.heuristic_tables(tree)
    c = dih[3 * {GLY: 0, PRO: 1}.get(a2, 2) + {GLY: 0, PRO: 1}.get(a3, 2)]
    L = 3.8
    for phi in (-2.5, -0.7, 0.4, 1.9):
        # beads 1..4 with dihedral phi about the z axis (phi > 0 when
        # (r12 x r23) . r34 > 0, as the reference signs it)
        pos = [[L, 0, 0], [0, 0, 0], [0, 0, L], [L * math.cos(phi), L * math.sin(phi), L]]
        ora = Oracle(p, chain_system(pos, [ALA, a2, a3, ALA]))
        _, V = ora.eval_forces(1 << KD)
        s, co = math.sin(phi), math.cos(phi)
        want = c[0] + c[1] * s + c[2] * co + c[3] * s * s + c[4] * co * co + c[5] * s * co
        assert V[0, KD] == pytest.approx(want, rel=1e-12, abs=1e-12)


@pytest.fixture(scope="module")
def heur_chain():
    rng = np.random.default_rng(41)
    pos = synth.saw_chain(40, rng)
    at = rng.integers(0, 20, size=40)
    at[[3, 9, 10, 17]] = GLY
    at[[4, 11, 12, 25]] = PRO
    return chain_system(pos, at)


@pytest.mark.parametrize("term", ["heur_ang", "heur_dih"])
def test_force_is_minus_gradient(term, heur_chain):
    k = _abi.TERM[term]
    p = heur_params(angles=term == "heur_ang", dihedrals=term == "heur_dih")
    ora = Oracle(p, heur_chain)
    rng = np.random.default_rng(5)
    r = heur_chain.r0 + rng.normal(scale=0.05, size=heur_chain.r0.shape)
    ora.set_positions(0, r)
    F, V = ora.eval_forces(1 << k)
    assert np.abs(F).max() > 0
    h = 1e-6
    for i in rng.choice(heur_chain.n_beads, size=12, replace=False):
        for a in range(3):
            rp = r.copy(); rp[a, i] += h
            ora.set_positions(0, rp)
            Vp = ora.eval_forces(1 << k)[1][0, k]
            rm = r.copy(); rm[a, i] -= h
            ora.set_positions(0, rm)
            Vm = ora.eval_forces(1 << k)[1][0, k]
            fd = -(Vp - Vm) / (2 * h)
            assert F[a, i] == pytest.approx(fd, rel=1e-5, abs=1e-6 * max(1.0, np.abs(F).max())), (term, i, a)
    assert np.abs(F.sum(axis=1)).max() < 1e-9 * max(1.0, np.abs(F).max())
    torque = np.cross(r.T, F.T).sum(axis=0)
    assert np.abs(torque).max() < 1e-8 * max(1.0, np.abs(F).max())


def test_native_value_takes_precedence(heur_chain):
    """An angle or dihedral with a native value stays with the native term;
    unflagged ones get no heuristic term."""
    p = heur_params()
    mask = (1 << KA) | (1 << KD)
    _, V_all = Oracle(p, heur_chain).eval_forces(mask)
    theta = heur_chain.nat_theta.copy()
    theta[5] = 1.9
    phi = heur_chain.nat_phi.copy()
    phi[7] = 0.5
    sys2 = dataclasses.replace(heur_chain, nat_theta=theta, nat_phi=phi)
    bits = heur_chain.heur_bonded.copy()
    bits[12] = 0
    sys3 = dataclasses.replace(heur_chain, heur_bonded=bits)
    for s, drop_a, drop_d in ((sys2, 5, 7), (sys3, 12, 12)):
        _, V = Oracle(p, s).eval_forces(mask)
        only = lambda start, nbead, kk: Oracle(p, dataclasses.replace(
            heur_chain, heur_bonded=np.where(np.arange(heur_chain.n_beads) == start,
                                             heur_chain.heur_bonded & (1 << kk), 0).astype(np.uint8)))
        Va = only(drop_a, 3, 0).eval_forces(1 << KA)[1][0, KA]
        Vd = only(drop_d, 4, 1).eval_forces(1 << KD)[1][0, KD]
        assert V[0, KA] == pytest.approx(V_all[0, KA] - Va, rel=1e-12, abs=1e-12)
        assert V[0, KD] == pytest.approx(V_all[0, KD] - Vd, rel=1e-12, abs=1e-12)


def test_terms_off_by_default():
    p, _ = P.from_yaml(None)
    assert p.heur_ang_enabled == 0 and p.heur_dih_enabled == 0
    p2, _ = P.from_yaml({"heurestic angles": {"enabled": True}})
    assert p2.heur_ang_enabled == 1 and p2.heur_ang_coeffs[8][0] == 15883.02041
