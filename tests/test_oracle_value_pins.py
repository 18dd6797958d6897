"""Closed-form value pins of the oracle's angle-dependent energies.

Finite differences, Newton's third law and the torque balance fix the
forces relative to the energy, not the energy itself: a consistent mistake
in an angle convention or a parameter table passes all of them.  Here each
energy is computed on hand-built geometries from its definition alone:

  * PID bundle (xmd/src/forces/pid/eval_forces.cpp:27-126,
    xmd/include/xmd/forces/pid/lambda.h:34-41, bundle.cpp:22-24):
    V = sum over bb+, bb-, ss of lambda(psi1) lambda(psi2) LJ(r), with the
    pseudo-improper angles psi from the textbook (IUPAC) dihedral of the
    construction and an asymmetric side-side table;
  * native angle (angle/native.cpp:29-33): V = k (theta - theta_n)^2 / 2
    for a bond angle theta built into the chain;
  * native dihedrals (dihedral/complex_native.cpp:26-36,
    simple_native.cpp:26-30): V = CDA (1 - cos D) + CDB (1 - cos 3D) and
    CDH D^2 / 2 with D = phi - phi_n for a quad twisted by phi;
  * chirality (chiral.cpp:27-30): V = e (C - C_n)^2 / 2 with C the
    determinant of the three bond vectors times the native factor;
  * relative Debye-Hueckel (es/relative.cpp:23-24) at r = s;
  * every bonded term vanishes, with its force, on the synthetic native
    chain (the natives are computed by numpy in xmd_amd/synth.py).

test_pins_reject_planted_mistakes builds the oracle with each mistake of
its ORACLE_MUTANT list (all of them keep F = -grad V) and checks that these
pins reject every one.
"""
import math
import os
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from xmd_amd import _abi, synth
from xmd_amd.params import AA_INDEX, ANGSTROM
from oracle.oracle import Oracle

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
BOND = 3.8  # A


def only(term, cid=1):
    p = synth.config_params(cid)
    synth.apply_terms(p, {k: False for k in synth._FLAG})
    synth.apply_terms(p, {term: True})
    return p


def one_replica(chains):
    """System of one open replica; chains = [(positions in A, residue names)]."""
    return synth.make_system([dict(chains=[(np.asarray(pos, float), [AA_INDEX[a] for a in aa])
                                           for pos, aa in chains],
                                   box_A=(0, 0, 0), temperature=0.35, seed=1)])


def iupac_dihedral(p0, p1, p2, p3):
    """Signed dihedral of p0-p1-p2-p3 (IUPAC sign), atan2 form; the
    arguments may carry leading batch axes."""
    b0, b1, b2 = p1 - p0, p2 - p1, p3 - p2
    n1, n2 = np.cross(b0, b1), np.cross(b1, b2)
    y = np.linalg.norm(b1, axis=-1) * np.sum(b0 * n2, axis=-1)
    return np.arctan2(y, np.sum(n1 * n2, axis=-1))


def lj(r, r_min, depth):
    s6 = (r_min / r) ** 6
    return float(np.clip(depth * (s6 * s6 - 2 * s6), -1e3, 1e3))


def sink_lj(r, r_min, r_max, depth):
    return lj(r, min(max(r, r_min), r_max), depth) if r > 0 else 0.0


def lam_cos(psi, psi0, alpha):
    """Cosine lambda and whether psi is in its support."""
    s = alpha * (psi - psi0)
    return 0.5 * math.cos(s) + 0.5, abs(s) < math.pi


def term_energy(make, params, system, term):
    """Energy of `term` and the forces, from `make(params, system)` (an
    Oracle or an Engine)."""
    ora = make(params, system)
    ora.rebuild_lists()
    F, V = ora.eval_forces()
    return V[0, _abi.TERM[term]], F


# ----------------------------------------------------------------- PID ---
def pid_side(cen, partner, psi_target, h=1.5):
    """prev and next of `cen` (bonds of 3.8 A, cen h A off the prev-next
    axis, the axis perpendicular to cen -> partner), cen's offset from the
    axis turned about the axis so that the pseudo-improper angle of (prev,
    cen, next; partner) is psi_target.

    The reference's angle (eval_forces.cpp:27-57) is that between the planes
    (prev, cen, next) and (prev, next, partner), with normals
    (next - cen) x (cen - prev) and (next - prev) x (partner - prev), made
    negative when (cen - next) . (second normal) < 0.  That is minus the
    IUPAC dihedral of (cen, prev, next, partner): the cis arrangement
    (partner on cen's side of the axis, in its plane) is 0, and a partner
    displaced against (cen - prev) x (next - cen) is positive."""
    a = math.sqrt(BOND ** 2 - h ** 2)
    u = (partner - cen) / np.linalg.norm(partner - cen)
    e = np.cross(u, [0.3, 0.5, 0.8])
    e /= np.linalg.norm(e)  # the prev -> next axis, perpendicular to u

    def build(ang):
        # cen's offset from the axis midpoint: -u turned about e by ang
        ang = np.asarray(ang)[..., None]
        y = -u * np.cos(ang) + np.cross(e, -u) * np.sin(ang)
        m = cen - h * y
        prev, nxt = m - a * e, m + a * e
        return prev, nxt, -iupac_dihedral(cen, prev, nxt, partner)

    lo, hi = -math.pi, math.pi
    for _ in range(3):  # grid, then two refinements around the best point
        ang = np.linspace(lo, hi, 4001)
        psi = build(ang)[2]
        k = int(np.argmin(np.abs((psi - psi_target + math.pi) % (2 * math.pi) - math.pi)))
        step = (hi - lo) / 4000
        lo, hi = ang[k] - step, ang[k] + step
    prev, nxt, psi = build(ang[k])
    assert abs(psi - psi_target) < 1e-4
    return prev, nxt, float(psi)


def pid_case(psi1, psi2, r, aa1="SER", aa2="ILE", tilt=(0.0, 0.0, 0.0)):
    c1 = np.zeros(3)
    c2 = np.array([r, 0.0, 0.0]) + np.asarray(tilt)
    p1, n1, ps1 = pid_side(c1, c2, psi1)
    p2, n2, ps2 = pid_side(c2, c1, psi2)
    system = one_replica([(np.stack([p1, c1, n1]), ["GLY", aa1, "GLY"]),
                          (np.stack([p2, c2, n2]), ["GLY", aa2, "GLY"])])
    return system, ps1, ps2, float(np.linalg.norm(c2 - c1))


def pid_params():
    """PID alone, with a side-side table that is not symmetric in the
    residue pair (the defaults' is)."""
    p = only("pid", cid=2)
    a, b = AA_INDEX["SER"], AA_INDEX["ILE"]
    p.ss_rmax[20 * a + b] = 7.0 * ANGSTROM
    p.ss_rmax[20 * b + a] = 7.6 * ANGSTROM
    p.ss_depth[20 * a + b] = 1.3
    p.ss_depth[20 * b + a] = 0.7
    return p


def pid_expected(p, psi1, psi2, r, aa1="SER", aa2="ILE"):
    ty = 20 * AA_INDEX[aa1] + AA_INDEX[aa2]  # bundle.cpp:22-24: i1's residue first
    ss_rmin = p.lj_bb_rmin if p.ss_sinking else p.ss_rmax[ty]
    V = 0.0
    for psi0, alpha, ljv in (
            (p.pid_bbp_psi0, p.pid_bbp_alpha, lj(r, p.pid_bbp_rmin, p.pid_bbp_depth)),
            (p.pid_bbm_psi0, p.pid_bbm_alpha, lj(r, p.pid_bbm_rmin, p.pid_bbm_depth)),
            (p.pid_ss_psi0, p.pid_ss_alpha, sink_lj(r, ss_rmin, p.ss_rmax[ty], p.ss_depth[ty]))):
        l1, in1 = lam_cos(psi1, psi0, alpha)
        l2, in2 = lam_cos(psi2, psi0, alpha)
        if in1 and in2:
            V += l1 * l2 * ljv
    return V


# psi pairs: bb+ and ss; bb- and ss; ss only (both signs); r beyond and
# inside the side-side flat bottom
PID_CASES = [(0.9, 1.2, 8.0), (1.3, 0.7, 6.3), (-1.3, -1.6, 8.0), (-1.1, -1.0, 5.8),
             (0.8, -1.2, 8.0), (2.0, 0.3, 7.3), (-2.5, 1.9, 6.0)]


def check_pid(make):
    p = pid_params()
    for psi1, psi2, r_A in PID_CASES:
        system, ps1, ps2, r = pid_case(psi1, psi2, r_A)
        V, _ = term_energy(make, p, system, "pid")
        expect = pid_expected(p, ps1, ps2, r * ANGSTROM)
        assert V == pytest.approx(expect, rel=1e-9, abs=1e-12), (psi1, psi2, r_A, V, expect)


def test_pid_construction_matches_its_angles():
    """The construction's psi by the reference's definition written as
    geometry: a partner rotated about the prev-next axis by +t away from
    cen's half-plane, towards (cen - prev) x (next - cen), has psi = t."""
    prev, cen, nxt = np.array([-3.0, 0, 0]), np.array([0, 1.5, 0]), np.array([3.0, 0, 0])
    # (cen - prev) x (next - cen) points along -z
    for t in (0.0, 0.4, 1.3, -0.7, 2.9):
        partner = np.array([0.5, 6 * math.cos(t), -6 * math.sin(t)])
        assert -float(iupac_dihedral(cen, prev, nxt, partner)) == pytest.approx(t, abs=1e-12)


def test_pid_bundle_closed_form():
    check_pid(Oracle)


def test_pid_covers_every_branch():
    """The cases exercise each lambda support and both LJ regimes."""
    p = pid_params()
    seen = set()
    for psi1, psi2, _ in PID_CASES:
        for name, psi0, alpha in (("bb+", p.pid_bbp_psi0, p.pid_bbp_alpha), ("bb-", p.pid_bbm_psi0, p.pid_bbm_alpha),
                                  ("ss", p.pid_ss_psi0, p.pid_ss_alpha)):
            if lam_cos(psi1, psi0, alpha)[1] and lam_cos(psi2, psi0, alpha)[1]:
                seen.add(name)
    assert seen == {"bb+", "bb-", "ss"}


# ------------------------------------------------------- native angle ---
ANGLE_CASES = [(100.0, 110.0), (150.0, 120.0), (62.0, 95.0)]


def check_angle(make):
    p = only("nat_ang")
    for th, th_n in ANGLE_CASES:
        t = math.radians(th)
        pos = np.array([[BOND, 0, 0], [0, 0, 0], [BOND * math.cos(t), BOND * math.sin(t), 0]])
        system = one_replica([(pos, ["ALA", "GLY", "VAL"])])
        system.nat_theta[0] = math.radians(th_n)
        V, _ = term_energy(make, p, system, "nat_ang")
        assert V == pytest.approx(0.5 * p.nat_ang_k * (t - math.radians(th_n)) ** 2, rel=1e-12), (th, th_n)


def test_native_angle_closed_form():
    check_angle(Oracle)


# ---------------------------------------------------- native dihedrals ---
# (phi, phi_n) in rad; D = phi - phi_n is not wrapped (simple_native.cpp:26)
DIH_CASES = [(1.0, -2.0), (-0.5, 0.7), (2.8, -2.9), (-2.0, -2.4)]


def dih_quad(phi):
    c, d = 1.3, math.sqrt(BOND ** 2 - 1.3 ** 2)
    pos = np.array([[-c, d, 0], [0, 0, 0], [BOND, 0, 0], [BOND + c, d * math.cos(phi), d * math.sin(phi)]])
    assert float(iupac_dihedral(*pos)) == pytest.approx(phi, abs=1e-12)
    return pos


def check_dihedrals(make):
    for term in ("cnd", "snd"):
        p = only(term)
        for phi, phi_n in DIH_CASES:
            system = one_replica([(dih_quad(phi), ["ALA", "GLY", "LEU", "PRO"])])
            system.nat_phi[0] = phi_n
            V, _ = term_energy(make, p, system, term)
            D = phi - phi_n
            if term == "cnd":
                expect = p.cnd_CDA * (1 - math.cos(D)) + p.cnd_CDB * (1 - math.cos(3 * D))
            else:
                expect = 0.5 * p.snd_CDH * D * D
            assert V == pytest.approx(expect, rel=1e-11), (term, phi, phi_n)


def test_native_dihedrals_closed_form():
    check_dihedrals(Oracle)


# ----------------------------------------------------------- chirality ---
def check_chirality(make):
    p = only("chir")
    rng = np.random.default_rng(3)
    for k in range(4):
        pos = np.cumsum(rng.normal(size=(4, 3)), axis=0) * 2.0
        system = one_replica([(pos, ["ALA", "GLY", "LEU", "PRO"])])
        b = pos[1:] - pos[:-1]                                  # r12, r23, r34 in A
        fac = (0.9 * BOND * ANGSTROM) ** -3
        c_nat = (-1) ** k * 0.37
        system.nat_chir[0], system.nat_chir_factor[0] = c_nat, fac
        V, _ = term_energy(make, p, system, "chiral")
        C = np.linalg.det(b * ANGSTROM) * fac
        assert V == pytest.approx(0.5 * p.chir_e * (C - c_nat) ** 2, rel=1e-11), k


def test_chirality_closed_form():
    check_chirality(Oracle)


# --------------------------------------------- relative Debye-Hueckel ---
def check_relative_dh(make):
    p = only("rel_es")
    s_A = p.rel_es_screen / ANGSTROM
    system = one_replica([(np.array([[0.0, 0, 0]]), ["ARG"]), (np.array([[s_A, 0, 0]]), ["ASP"])])
    V, F = term_energy(make, p, system, "rel_es")
    s = p.rel_es_screen
    expect = (+1) * (-1) * math.exp(-1.0) / (4 * math.pi * p.rel_es_factor * s * s)
    assert V == pytest.approx(expect, rel=1e-12)
    # dV/dr = -V (1/s + 2/r) at r = s; the force on bead 0 is dV/dr along +x
    assert F[0, 0] == pytest.approx(-expect * 3.0 / s, rel=1e-12)
    assert F[0, 0] > 0  # opposite charges attract


def test_relative_debye_hueckel_at_screening_length():
    check_relative_dh(Oracle)


# ------------------------------------------------- natives are minima ---
def check_natives_vanish(make):
    system = synth.config_system(1, chain_len=60)
    for term in ("tether", "chir", "nat_ang", "cnd", "snd"):
        p = only(term)
        V, F = term_energy(make, p, system, {"chir": "chiral"}.get(term, term))
        assert abs(V) < 1e-18 and np.abs(F).max() < 1e-9, (term, V, np.abs(F).max())


def test_bonded_terms_vanish_on_native_chain():
    check_natives_vanish(Oracle)


# ------------------------------------------------- planted mistakes ---
CHECKS = [check_pid, check_angle, check_dihedrals, check_chirality, check_relative_dh, check_natives_vanish]
MUTANTS = {1: "PID bb+/bb- LJ swapped", 2: "PID side-side index transposed", 3: "PID psi sign flipped",
           4: "native angle supplement", 5: "native dihedral sign flipped", 6: "chirality sign flipped",
           7: "relative DH as 1/r"}


def _build_mutant(tmp, k):
    out = os.path.join(tmp, f"liboracle_m{k}.so")
    cmd = ["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared",
           f"-DORACLE_MUTANT={k}", "-o", out, os.path.join(ORACLE_DIR, "xmd_oracle.cpp")]
    subprocess.run(cmd, check=True, capture_output=True)
    return out


def _rejected(path):
    """Names of the checks that fail on the oracle at `path`."""
    failed = []
    for chk in CHECKS:
        try:
            chk(lambda p, s: Oracle(p, s, lib_path=path))
        except AssertionError:
            failed.append(chk.__name__)
    return failed


def test_pins_reject_planted_mistakes():
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    tmp = tempfile.mkdtemp(prefix="oracle_mutants_")
    try:
        with ThreadPoolExecutor(max_workers=min(7, os.cpu_count() or 1)) as ex:
            paths = dict(zip(MUTANTS, ex.map(lambda k: _build_mutant(tmp, k), MUTANTS)))
        missed = {MUTANTS[k]: _rejected(p) for k, p in paths.items()}
        assert all(missed.values()), {k: v for k, v in missed.items() if not v}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


@pytest.mark.gpu
@pytest.mark.parametrize("check", CHECKS, ids=lambda c: c.__name__)
def test_engine_closed_forms(check):
    """The same closed forms on the FP64 engine (device libm within an ulp
    or two of glibc, inside the checks' tolerances)."""
    from xmd_amd.engine import Engine
    check(Engine)
