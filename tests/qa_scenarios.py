"""Hand-built QA geometries whose outcome follows from geometry and the
default parameters alone (no oracle formulas).

Two 3-bead chains GLY-LEU-GLY face each other.  Middle beads have
n = unit(v2 - v1) and h = unit(v2 x v1) (qa/precompute_nh.cpp:17-19);
terminal GLY beads have n = h = 0 and GLY has r_max = 0 against every
residue, so the only possible contact is between the two LEU beads.

  chain A: (-3, 1, 0), (0, 0, 0), (3, 1, 0)   -> n1 = +y, h1 = -z
  side-side: chain B mirrored below at y = -d -> n2 = -y, u = -y
             n1.u = -1 < 0.5, -n2.u = -1 < 0.5, h.u = 0  => SS(LEU, LEU)
  back-back: chain B = chain A shifted by -d z -> u = -z, h1.u = 1,
             h2.u = 1, h1.h2 = 1 => BB; n1.u = 0 < 0.5 so side1 is set too
Distances in A.
"""
import numpy as np

from xmd_amd import synth

GLY, LEU = 7, 10
LEU_LIMITS = (2, 7, 2, 4)      # back, side_all, side_polar, side_hydrophobic
SS_LEU_LEU = 4 + LEU + 20 * LEU
R_MAX_LEU_LEU = 8.93237        # "lj force variants" ss r_max table
BB_R_MIN = 5.0
FACTOR = 1.5 * 2.0 ** (-1.0 / 6.0)


def geometry(kind, d):
    a = np.array([[-3.0, 1, 0], [0, 0, 0], [3, 1, 0]])
    if kind == "ss":
        b = np.array([[-3.0, -d - 1, 0], [0, -d, 0], [3, -d - 1, 0]])
    else:
        b = a + np.array([0, 0, -d])
    types = [GLY, LEU, GLY]
    return [(a, types), (b, types)]


def qa_params():
    p = synth.config_params(2)
    synth.apply_terms(p, {k: False for k in ("go", "pauli", "tether", "chir", "nat_ang", "cnd", "snd",
                                             "const_es", "rel_es", "pid")})
    synth.apply_terms(p, {"qa": True})
    return p


def qa_system(kind, d):
    return synth.make_system([dict(chains=geometry(kind, d), box_A=(0, 0, 0), temperature=0.3, seed=3)])


def expected_contact(kind):
    """(type, diff1, diff2) of the LEU-LEU contact."""
    if kind == "ss":
        # side only; partner LEU is hydrophobic
        return SS_LEU_LEU, (0, 1, 0, 1), (0, 1, 0, 1)
    # back-back with side1 (n1.u = 0) and side2 (-n2.u = 0) also true
    return 1, (1, 1, 0, 1), (1, 1, 0, 1)


def run_timeline(sysobj, kind, d):
    """Drive one contact through FORMING -> BREAKING -> removal with frozen
    positions; returns a list of (t, n_contacts, status, V_qa, sync_of_bead1)."""
    from xmd_amd import _abi
    out = []
    y, _, rng, _ = sysobj.get_state(0)
    for t in (0.0, 2.5, 5.0, 10.0, 15.0, 20.0, 20.005):
        sysobj.set_state(0, y, rng, t)
        sysobj.qa_update()
        cont, sync = sysobj.get_qa_state(0)
        _, V = sysobj.eval_forces(1 << _abi.TERM["qa"])
        status = int(cont[0]["status"]) if len(cont) else -1
        out.append((t, len(cont), status, float(V[0, _abi.TERM["qa"]]), tuple(int(x) for x in sync[1])))
    return out


def expected_timeline(kind, d):
    """Closed form of run_timeline with frozen positions.

    saturation = min(t - t0, 10)/10 (phase duration 10 tau, exactly 1 at
    t - t0 = 10); a FORMING contact at saturation 1 turns BREAKING (ref = t)
    when factor * r_min / r < 1, r_min of the contact's LJ (the sinking ss
    LJ uses the bb r_min, 5 A); a BREAKING contact (saturation 1 - ramp) is
    removed when that reaches 0, restoring the budgets, and its pair forms
    again at the next update.  The sinking LJ is flat (V = -1) between r_min
    and r_max; BB uses the plain LJ of depth 1 at r_min 5 A."""
    _, d1, _ = expected_contact(kind)
    if kind == "ss":
        assert BB_R_MIN < d < R_MAX_LEU_LEU
        V_full = -1.0
    else:
        s = BB_R_MIN / d
        V_full = s ** 12 - 2 * s ** 6
    breaks = FACTOR * BB_R_MIN / d < 1.0
    lim = tuple(LEU_LIMITS)
    formed = tuple(int(x) for x in np.array(LEU_LIMITS) - np.array(d1))
    rows = [(0.0, 1, 0, 0.0, formed), (2.5, 1, 0, 0.25 * V_full, formed), (5.0, 1, 0, 0.5 * V_full, formed)]
    if not breaks:
        rows += [(t, 1, 0, V_full, formed) for t in (10.0, 15.0, 20.0, 20.005)]
    else:
        rows += [(10.0, 1, 1, V_full, formed), (15.0, 1, 1, 0.5 * V_full, formed), (20.0, 0, -1, 0.0, lim),
                 (20.005, 1, 0, 0.0, formed)]
    return rows


CASES = [("ss", 6.0), ("ss", 7.0), ("bb", 4.8)]


def check_timeline(got, want):
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g[0] == w[0]
        assert g[1:3] == w[1:3], (g, w)
        assert abs(g[3] - w[3]) <= 1e-12 * max(1.0, abs(w[3])), (g, w)
        assert g[4] == w[4], (g, w)
