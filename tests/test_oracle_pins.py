"""Pins of the CPU oracle that do not rely on the oracle's own formulas.

The reference has no tests or golden vectors for this path, so each oracle
function is checked against something the mathematics fixes: closed forms,
finite-difference gradients, Newton's third law, rotation invariance, an
independent numpy brute force of the list criterion, glibc called directly
for the noise transform, and integrator invariants.
"""
import ctypes as C
import math

import numpy as np
import pytest

from xmd_amd import _abi, synth
from xmd_amd.params import ANGSTROM
from oracle import oracle as O
from oracle.oracle import Oracle

LIBM = C.CDLL("libm.so.6")
for fn in ("logf", "sinf", "cosf", "sqrtf"):
    getattr(LIBM, fn).restype = C.c_float
    getattr(LIBM, fn).argtypes = [C.c_float]


# ------------------------------------------------------------------ RNG ---
MASK = (1 << 64) - 1


def xorshift_py(s):
    out = (s * 0xD989BCACC137DCD5) & MASK
    s ^= s >> 11
    s ^= (s << 31) & MASK
    s ^= s >> 18
    return out, s


def test_rng_stream_matches_closed_form():
    s = 0x5EED0001
    so = s
    for _ in range(1000):
        v, s = xorshift_py(s)
        vo, so = O.rng_next(so)
        assert v == vo and s == so


def _f32(x):
    return float(np.float32(x))


def normals_py(state, n):
    """Box-Muller of rand_gen with glibc float libm called directly."""
    out = np.zeros((n, 3), dtype=np.float32)
    s = state
    two_pi = _f32(2.0 * math.pi)
    inv = _f32(1.0 / 4294967296.0)

    def uni():
        nonlocal s
        v, s = xorshift_py(s)
        return _f32(_f32(v >> 32) * inv)

    for k in range(n):
        vals = []
        for kind in ("x2", "one"):
            r1, r2 = uni(), uni()
            lg = LIBM.logf(r1)
            r = LIBM.sqrtf(_f32(-2.0 * lg))
            if math.isnan(r):
                r = 0.0
            t = _f32(two_pi * r2)
            vals.append(_f32(r * LIBM.cosf(t)))
            if kind == "x2":
                vals.append(_f32(r * LIBM.sinf(t)))
        out[k] = vals
    return out, s


def test_noise_transform_matches_glibc():
    for state in (1, 0x5EED0000, 987654321987):
        a, sa = normals_py(state, 300)
        b, sb = O.rand_normals(state, 300)
        assert sa == sb
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_noise_moments():
    z, _ = O.rand_normals(42, 200000)
    z = z.astype(np.float64)
    assert abs(z.mean()) < 0.01
    assert abs(z.var() - 1.0) < 0.01
    c = np.corrcoef(z.T)
    assert np.abs(c - np.eye(3)).max() < 0.01


# ----------------------------------------------------------- primitives ---
def test_lj_closed_forms():
    V, dV = O.prim(0, 1.5, 2.0, 1 / 2.0)          # at r_min
    assert V == pytest.approx(-1.5) and abs(dV) < 1e-12
    V, dV = O.prim(0, 1.0, 2.0, 1 / 4.0)          # r = 2 r_min
    s6 = 2.0 ** -6
    assert V == pytest.approx(s6 * s6 - 2 * s6) and dV == pytest.approx(12 / 4.0 * (s6 - s6 * s6))
    V, dV = O.prim(0, 1.0, 2.0, 1 / 0.5)          # deep core: clamped
    assert V == 1e3 and dV == -1e3


def test_shifted_lj_and_harmonic_zeros():
    V, dV = O.prim(1, 1.0, 0.8, 1 / 0.8)
    assert abs(V) < 1e-12 and abs(dV) < 1e-12
    V, dV = O.prim(3, 100.0, 0.0, 0.76, 0.76)
    assert V == 0 and dV == 0
    V, dV = O.prim(3, 100.0, 2.0, 0.76, 0.86)
    d = 0.1
    assert V == pytest.approx(0.5 * d * d * (100 + 2 * d * d)) and dV == pytest.approx(d * (100 + 4 * d * d))


def test_sink_lj_flat_bottom():
    for r in (1.2, 1.5, 1.9):
        V, dV = O.prim(2, 1.0, 1.0, 2.0, r, 1 / r)
        assert dV == 0.0
    V, dV = O.prim(2, 1.0, 1.0, 1.0, 1.3, 1 / 1.3)  # r_min == r_max -> plain LJ
    V2, dV2 = O.prim(0, 1.0, 1.0, 1 / 1.3)
    assert V == pytest.approx(V2) and dV == pytest.approx(dV2)


@pytest.mark.parametrize("kind", [4, 5])
def test_lambda_value_and_derivative(kind):
    psi0, alpha = 0.3, 2.5
    # 16 points, none at psi0 where the algebraic form (a function of |s|) has a cusp
    for psi in np.linspace(psi0 - 1.2, psi0 + 1.2, 16):
        v, d = O.prim(kind, psi0, alpha, psi)
        h = 1e-6
        vp, _ = O.prim(kind, psi0, alpha, psi + h)
        vm, _ = O.prim(kind, psi0, alpha, psi - h)
        assert d == pytest.approx((vp - vm) / (2 * h), rel=1e-6, abs=1e-8)
    if kind == 4:
        assert O.prim(4, psi0, alpha, psi0)[0] == pytest.approx(1.0)
        assert O.prim(4, psi0, alpha, psi0 + math.pi / alpha)[0] == pytest.approx(0.0, abs=1e-12)


# -------------------------------------------------- per-term gradients ---
def _single_term(cid, term, **kw):
    p = synth.config_params(cid)
    for name in _abi.TERMS:
        synth.apply_terms(p, {synth_key(name): False for name in _abi.TERMS if synth_key(name)})
    synth.apply_terms(p, {synth_key(term): True})
    return p


def synth_key(term):
    return {"tether": "tether", "chiral": "chir", "nat_ang": "nat_ang", "cnd": "cnd", "snd": "snd",
            "pauli": "pauli", "go": "go", "const_es": "const_es", "rel_es": "rel_es", "qa": "qa",
            "pid": "pid", "nat_ss": "nat_ss", "solid_wall": "solid_wall", "lj_wall": "lj_wall",
            "vel_afm": "vel_afm", "force_afm": "force_afm", "heur_ang": "heur_ang", "heur_dih": "heur_dih"}[term]


@pytest.fixture(scope="module")
def small_open():
    return synth.config_system(1, chain_len=40)


@pytest.fixture(scope="module")
def small_pbc():
    return synth.config_system(2, chains=3, chain_len=40)


def _compressed(system, factor):
    """Shrink intra-chain geometry so that repulsive and PID terms are active."""
    r = system.r0.copy()
    return r * factor + (1 - factor) * r.mean(axis=1, keepdims=True)


FD_TERMS = ["tether", "chiral", "nat_ang", "cnd", "snd", "pauli", "go", "const_es", "rel_es", "pid"]


@pytest.mark.parametrize("term", FD_TERMS)
def test_force_is_minus_gradient(term, small_open, small_pbc):
    system = small_pbc if term == "pid" else small_open
    cid = 2 if term == "pid" else 1
    p = _single_term(cid, term)
    ora = Oracle(p, system)
    rng = np.random.default_rng(7)
    # non-native pairs are >= 7.5 A apart natively: Pauli needs a stronger squeeze
    r = _compressed(system, 0.5 if term == "pauli" else 0.8) + rng.normal(scale=0.05, size=system.r0.shape)
    ora.set_positions(0, r)
    ora.rebuild_lists()
    F, V = ora.eval_forces()
    k = _abi.TERM[term]
    assert np.abs(F).max() > 0, "term inactive in the test geometry"
    h = 1e-6
    n = system.n_beads
    idx = rng.choice(n, size=12, replace=False)
    for i in idx:
        for a in range(3):
            rp = r.copy(); rp[a, i] += h
            ora.set_positions(0, rp); ora.rebuild_lists()
            _, Vp = ora.eval_forces()
            rm = r.copy(); rm[a, i] -= h
            ora.set_positions(0, rm); ora.rebuild_lists()
            _, Vm = ora.eval_forces()
            fd = -(Vp[0, k] - Vm[0, k]) / (2 * h)
            assert F[a, i] == pytest.approx(fd, rel=1e-5, abs=1e-6 * max(1.0, np.abs(F).max())), (term, i, a)


@pytest.mark.parametrize("term", FD_TERMS)
def test_newton_third_law_and_torque(term, small_open, small_pbc):
    system = small_pbc if term == "pid" else small_open
    cid = 2 if term == "pid" else 1
    ora = Oracle(_single_term(cid, term), system)
    r = _compressed(system, 0.5 if term == "pauli" else 0.8)
    ora.set_positions(0, r)
    F, _ = ora.eval_forces()
    scale = np.abs(F).max()
    assert np.abs(F.sum(axis=1)).max() < 1e-9 * max(1, scale)
    if term != "pid":  # open box: total torque vanishes
        torque = np.cross(r.T, F.T).sum(axis=0)
        assert np.abs(torque).max() < 1e-8 * max(1, scale)


def test_debye_hueckel_value_at_screening_length():
    """Two charged beads at r = s: V = Vf q1 q2 e^-1 / s (constant
    permittivity, decision D9)."""
    system = synth.make_system([dict(chains=[(np.array([[0, 0, 0], [10.0, 0, 0]]), [1, 3])],
                                     box_A=(0, 0, 0), temperature=0.3, seed=1)])
    p = _single_term(1, "const_es")
    p.max_cutoff_override = 30 * ANGSTROM
    # chain neighbours are excluded from the list; use two chains instead
    system = synth.make_system([dict(chains=[(np.array([[0.0, 0, 0]]), [1]), (np.array([[10.0, 0, 0]]), [3])],
                                     box_A=(0, 0, 0), temperature=0.3, seed=1)])
    ora = Oracle(p, system)
    F, V = ora.eval_forces()
    s = 10 * ANGSTROM
    Vf = 1 / (4 * math.pi * p.const_es_permittivity)
    expect = Vf * (1.0 * -1.0) * math.exp(-1.0) / s
    assert V[0, _abi.TERM["const_es"]] == pytest.approx(expect, rel=1e-12)
    # force along the axis: -dV/dr = V (1/s + 1/r) pushes... sign check
    assert F[0, 0] == pytest.approx(-expect * (1 / s + 1 / s), rel=1e-12)


# ------------------------------------------------------------- lists ---
def brute_force_lists(system, p, rep=0):
    """Independent numpy restatement of the list criterion."""
    sl = system.replica_slice(rep)
    r = system.r0[:, sl].T
    n = r.shape[0]
    L = system.box[rep]
    Linv = np.where(L != 0, 1.0 / np.where(L != 0, L, 1), 0.0)
    chain = np.searchsorted(system.chain_offset, np.arange(sl.start, sl.stop), side="right") - 1
    # max cutoff as the reference aggregates it (no override in this test)
    ss = np.array(p.ss_rmax)
    cands = []
    if p.pauli_enabled:
        cands.append(p.pauli_r_excl)
    if p.go_enabled:
        cands.append((2.0 * system.nat_dist).max())
    es_cut = 2.0 * p.const_es_screen
    if p.const_es_enabled:
        cands.append(es_cut)
    mfmd = max(p.lj_bb_rmin, p.lj_bs_rmin, ss.max())
    if p.qa_enabled:
        cands.append(mfmd)
    pid_cut = max(2 * p.lj_bb_rmin, 2 * p.lj_bs_rmin, (2 * ss).max())
    if p.pid_enabled:
        cands.append(pid_cut)
    mc = max(cands)
    pad = p.pad_factor * mc
    req = mc + pad
    native = set(zip(system.nat_i1.tolist(), system.nat_i2.tolist()))
    out = {k: [] for k in ("nl_all", "nl_native", "pauli", "es", "qa_free")}
    q = np.array(p.aa_charge)
    for i in range(n):
        d = r[i + 1:] - r[i]
        d = d - L * np.rint(d * Linv)
        dist = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
        for jj in np.flatnonzero(dist < req):
            j = i + 1 + jj
            if chain[i] == chain[j] and j - i < 3:
                continue
            out["nl_all"].append((i, j))
            nat = (i + sl.start, j + sl.start) in native
            inv = 1.0 / dist[jj]
            if nat:
                out["nl_native"].append((i, j))
            if not nat and p.pauli_enabled and inv > 1.0 / (p.pauli_r_excl + pad):
                out["pauli"].append((i, j))
            if p.const_es_enabled and inv > 1.0 / (es_cut + pad) and q[system.atype[i]] * q[system.atype[j]] != 0:
                out["es"].append((i, j))
            if not nat and p.qa_enabled and inv > 1.0 / (mfmd + pad):
                out["qa_free"].append((i, j))
    return out


def odd_box_system(kind):
    """4 x 60 beads of the C2 density in a non-cubic periodic box
    (100 x 134 x 180 A) or a slab periodic in x and y only (z open)."""
    import dataclasses
    system = synth.config_system(2, chains=4, chain_len=60)
    L = system.box[0, 0]
    box = np.array([[100.0 * synth.ANGSTROM, L, 180.0 * synth.ANGSTROM]]) if kind == "ortho" \
        else np.array([[L, L, 0.0]])
    return dataclasses.replace(system, box=box)


@pytest.mark.parametrize("cid", [1, 2, "ortho", "slab"])
def test_lists_match_brute_force(cid):
    if cid in ("ortho", "slab"):
        system = odd_box_system(cid)
    else:
        system = synth.config_system(1, chain_len=150) if cid == 1 else synth.config_system(2, chains=4, chain_len=60)
    p = synth.config_params(2 if cid in ("ortho", "slab") else cid)
    ora = Oracle(p, system)
    ora.rebuild_lists()
    ref = brute_force_lists(system, p)
    for name, pairs in ref.items():
        a, b = ora.get_list(0, name)
        got = list(zip(a.tolist(), b.tolist()))
        assert got == pairs, name


# -------------------------------------------------------- integrator ---
def _free_beads(n, T=0.35, gamma=2.0):
    chains = [(np.array([[20.0 * k, 0, 0]]), [0]) for k in range(n)]
    system = synth.make_system([dict(chains=chains, box_A=(0, 0, 0), temperature=T, seed=11)])
    p = synth.config_params(0)
    synth.apply_terms(p, {k: False for k in ("go", "pauli", "tether", "chir", "nat_ang", "cnd", "snd",
                                             "const_es", "rel_es", "qa", "pid")})
    p.gamma_factor = gamma
    return p, system


def test_free_flight_is_exact():
    p, system = _free_beads(4, T=0.0, gamma=0.0)
    ora = Oracle(p, system)
    y, _, rng, t = ora.get_state(0)
    y = y.copy()
    v = np.array([0.3, -0.2, 0.1])
    y[1] = (v * p.dt)[:, None] * np.ones((3, 4))
    ora.set_state(0, y, rng, 0.0)
    ora.step(100)
    y2, vel, _, t2 = ora.get_state(0)
    assert t2 == pytest.approx(100 * p.dt)
    np.testing.assert_allclose(y2[0], y[0] + 100 * y[1], rtol=0, atol=1e-12)
    np.testing.assert_allclose(vel, v[:, None] * np.ones((3, 4)), rtol=1e-12)


def test_friction_decay_matches_exponential():
    p, system = _free_beads(2, T=0.0, gamma=2.0)
    ora = Oracle(p, system)
    y, _, rng, _ = ora.get_state(0)
    y = y.copy()
    # scaled derivatives y_k = dt^k/k! x^(k) of v(t) = v0 exp(-gamma t)
    v0, g, dt = 0.5, 2.0, p.dt
    for k in range(1, 6):
        y[k] = dt ** k / math.factorial(k) * v0 * (-g) ** (k - 1)
    ora.set_state(0, y, rng, 0.0)
    ora.step(400)
    _, vel, _, t = ora.get_state(0)
    assert vel[0, 0] == pytest.approx(v0 * math.exp(-g * t), rel=1e-7)


def gear_stationary_velocity_variance(gamma, T, dt):
    """Exact stationary Var(v) of the Gear-5 recurrence of a free bead driven
    by the Langevin noise: y <- P (I - c e^T) y + P c (sigma dt^2/2) xi with
    sigma^2 = 2 gamma T (decision D13: no dt^-1/2), solved as a discrete
    Lyapunov equation.  P is the Pascal predictor on (y1..y5), c the
    textbook Gear coefficients, e^T y = y2 + gamma dt/2 y1."""
    from scipy.linalg import solve_discrete_lyapunov
    c = np.array([251 / 360, 1.0, 11 / 18, 1 / 6, 1 / 60])
    e = np.array([gamma * dt / 2, 1, 0, 0, 0])
    P = np.array([[math.comb(j + 1, i + 1) if j >= i else 0 for j in range(5)] for i in range(5)], dtype=float)
    A = P @ (np.eye(5) - np.outer(c, e))
    B = P @ c * (math.sqrt(2 * gamma * T) * dt * dt / 2)
    S = solve_discrete_lyapunov(A, np.outer(B, B))
    return S[0, 0] / dt ** 2


def test_stationary_velocity_variance():
    p, system = _free_beads(2000, T=0.4, gamma=2.0)
    ora = Oracle(p, system)
    ora.step(1500)
    acc = []
    for _ in range(20):
        ora.step(25)
        _, vel, _, _ = ora.get_state(0)
        acc.append((vel ** 2).mean())
    assert np.mean(acc) == pytest.approx(gear_stationary_velocity_variance(2.0, 0.4, p.dt), rel=0.03)


# ------------------------------------------------------ disulfide bonds ---
def test_nat_ss_bonds(small_open):
    """Native disulfide bonds (native_ss.cpp): V = H1 (r - r0)^2 / 2 summed
    over the bonds within 2 r0 + pad (D23), F = -grad V, and the bonded
    pairs sit in the list's native_ssbonds only, outside every pair term
    (D24)."""
    import dataclasses
    base = small_open
    n = base.n_beads
    r = base.r0.copy()
    # three bonds: two within reach, one pulled far apart
    pairs = [(2, 20), (5, 30), (10, 39)]
    ctr = r[:, 20].copy()
    r[:, 2] = ctr + np.array([1.1, 0.3, -0.2])  # ~5.8 A from bead 20
    r[:, 30] = r[:, 5] + np.array([0.0, 1.4, 0.5])
    r[:, 39] = r[:, 10] + np.array([6.0, 0.0, 0.0])  # 30 A: beyond 2 r0 + pad
    system = dataclasses.replace(base, ss_i1=np.array([a for a, _ in pairs], dtype=np.int32),
                                 ss_i2=np.array([b for _, b in pairs], dtype=np.int32))
    p = _single_term(1, "nat_ss")
    p.pauli_enabled = 1  # Pauli must not see the bonded pairs
    ora = Oracle(p, system)
    ora.set_positions(0, r)
    ora.rebuild_lists()
    cur = set(zip(*ora.get_list(0, "nat_ss")))
    assert cur == {(2, 20), (5, 30)}
    in_nl = set(zip(*ora.get_list(0, "nl_nat_ss")))
    assert in_nl <= set(pairs) and (2, 20) in in_nl
    for name in ("nl_native", "nl_non_native", "pauli"):
        assert not (set(zip(*ora.get_list(0, name))) & set(pairs)), name
    F, V = ora.eval_forces(1 << _abi.TERM["nat_ss"])
    H1, r0 = p.nat_ss_H1, p.nat_ss_equil_dist
    want = sum(0.5 * H1 * (np.linalg.norm(r[:, j] - r[:, i]) - r0) ** 2 for i, j in [(2, 20), (5, 30)])
    assert V[0, _abi.TERM["nat_ss"]] == pytest.approx(want, rel=1e-13)
    h = 1e-6
    for i in (2, 20, 5, 30):
        for a in range(3):
            rp = r.copy(); rp[a, i] += h
            ora.set_positions(0, rp); ora.rebuild_lists()
            _, Vp = ora.eval_forces(1 << _abi.TERM["nat_ss"])
            rm = r.copy(); rm[a, i] -= h
            ora.set_positions(0, rm); ora.rebuild_lists()
            _, Vm = ora.eval_forces(1 << _abi.TERM["nat_ss"])
            fd = -(Vp[0, _abi.TERM["nat_ss"]] - Vm[0, _abi.TERM["nat_ss"]]) / (2 * h)
            assert F[a, i] == pytest.approx(fd, rel=1e-6, abs=1e-8)
    assert np.abs(F[:, 39]).max() == 0 and np.abs(F[:, 10]).max() == 0


# ------------------------------------------------------- walls and AFM ---
def walls_afm_system(base):
    """The chain between a solid wall under it (normal +z) and an LJ
    attractive wall beside it (normal +x), a velocity tip on bead 0 and a
    force tip on the last bead."""
    import dataclasses
    r = base.r0
    zmin, xmin = r[2].min(), r[0].min()
    solid = np.array([[0.0, 0.0, zmin - 0.6, 0.0, 0.0, 1.0]])
    lj = np.array([[xmin - 0.5, 0.0, 0.0, 1.0, 0.0, 0.0]])
    n = base.n_beads
    return dataclasses.replace(base, solid_walls=solid, lj_walls=lj,
                               vel_afm_idx=np.array([0], dtype=np.int32),
                               vel_afm_orig=np.array([[np.nan] * 3]), vel_afm_v=np.array([[0.0, 0.0, 0.05]]),
                               force_afm_idx=np.array([n - 1], dtype=np.int32),
                               force_afm_F=np.array([[0.3, -0.2, 0.1]]))


def walls_afm_params(cid):
    p = _single_term(cid, "tether")
    p.solid_walls_enabled = p.lj_walls_enabled = p.vel_afm_enabled = p.force_afm_enabled = 1
    return p


def test_solid_wall_closed_form_and_gradient(small_open):
    system = walls_afm_system(small_open)
    p = _single_term(1, "solid_wall")
    ora = Oracle(p, system)
    r = system.r0
    k = _abi.TERM["solid_wall"]
    F, V = ora.eval_forces(1 << k)
    x = r[2] - system.solid_walls[0, 2]
    near = x < p.solid_walls_cutoff
    assert near.sum() > 0
    assert V[0, k] == pytest.approx((p.solid_walls_depth / (9 * x[near] ** 9)).sum(), rel=1e-13)
    # pushes along +z only, -dV/dx = eps / x^10
    assert np.allclose(F[2, near], p.solid_walls_depth / x[near] ** 10, rtol=1e-13)
    assert np.abs(F[:2]).max() == 0 and np.abs(F[2, ~near]).max() == 0
    h = 1e-6
    i = int(np.argmin(x))
    rp, rm = r.copy(), r.copy()
    rp[2, i] += h
    rm[2, i] -= h
    ora.set_positions(0, rp); Vp = ora.eval_forces(1 << k)[1][0, k]
    ora.set_positions(0, rm); Vm = ora.eval_forces(1 << k)[1][0, k]
    assert F[2, i] == pytest.approx(-(Vp - Vm) / (2 * h), rel=1e-6)


def test_lj_wall_timeline(small_open):
    """Beads within wall min dist of the plane join it at the projection
    of their position; the LJ toward the joint ramps in over one phase
    duration; a joint stretched beyond factor * min dist breaks and ramps
    out."""
    system = walls_afm_system(small_open)
    p = _single_term(1, "lj_wall")
    p.dt = 0.005
    p.lj_walls_phase_duration = 0.05  # 10 steps
    ora = Oracle(p, system)
    st, jo, rt = ora.get_wall_state(0)
    assert np.all(st == 0)
    ora.step(1)
    st, jo, rt = ora.get_wall_state(0)
    r = system.r0
    x = r[0] - system.lj_walls[0, 0]
    joined = np.abs(x) < p.lj_walls_min_dist
    assert joined.sum() > 0 and np.array_equal(st == 1, joined)
    assert np.allclose(jo[joined, 0], system.lj_walls[0, 0]) and np.allclose(jo[joined, 1:], r[1:, joined].T)
    assert np.all(rt[joined] == 0.0)
    # the force toward the joint ramps with t / phase duration
    k = _abi.TERM["lj_wall"]
    ora.step(4)
    y, _, _, t = ora.get_state(0)
    _, V = ora.eval_forces(1 << k)
    i = np.flatnonzero(joined)
    d = np.linalg.norm(jo[i] - y[0][:, i].T, axis=1)
    rmin = p.lj_walls_min_dist
    lj = np.clip((rmin / d) ** 12 - 2 * (rmin / d) ** 6, -1e3, 1e3)  # lj.h clamps V
    assert 0 < t < 0.05
    assert V[0, k] == pytest.approx((lj * min(t, 0.05) / 0.05).sum(), rel=1e-12)
    # pull one joined bead far from its joint: breaks once saturated, then frees
    ora.step(10)
    y, _, _, t = ora.get_state(0)
    b = i[0]
    y[0][0, b] += 3 * rmin * p.lj_walls_breaking_factor
    ora.set_state(0, y, ora.get_state(0)[2], t)
    ora.step(1)
    st2, _, rt2 = ora.get_wall_state(0)
    assert st2[b] == 2 and rt2[b] == pytest.approx(t)
    ora.step(12)
    assert ora.get_wall_state(0)[0][b] in (0, 1)  # freed (and maybe joined again)


def test_afm_forces(small_open):
    system = walls_afm_system(small_open)
    p = _single_term(1, "vel_afm")
    p.force_afm_enabled = 1
    ora = Oracle(p, system)
    F, V = ora.eval_forces()
    # force tip: the constant force; velocity tip at t = 0 sits on its bead
    assert np.allclose(F[:, -1], [0.3, -0.2, 0.1], rtol=0, atol=0)
    assert np.abs(F[:, 0]).max() == 0
    y, _, rng, _ = ora.get_state(0)
    ora.set_state(0, y, rng, 2.0)  # the tip has moved 0.1 along +z
    F, V = ora.eval_forces()
    H1 = p.vel_afm_H1
    assert np.allclose(F[:, 0], [0.0, 0.0, H1 * 0.1], rtol=1e-12, atol=1e-15)
    assert V[0, _abi.TERM["vel_afm"]] == 0 and V[0, _abi.TERM["force_afm"]] == 0


def test_parallel_list_build_equals_serial():
    """The oracle's threaded list build (its rebuild entry point, from 4096
    beads) gives the serial legacy_update's lists in the same order."""
    system = synth.config_system(2, chains=20, chain_len=250)
    p = synth.config_params(2)
    a, b = Oracle(p, system), Oracle(p, system)
    a.rebuild_lists()  # threaded
    b.eval_forces()    # first evaluation: serial build
    for name in _abi.LISTS:
        x1, x2 = a.get_list(0, name)
        y1, y2 = b.get_list(0, name)
        assert np.array_equal(x1, y1) and np.array_equal(x2, y2), name
    assert len(a.get_list(0, "nl_all")[0]) > 10000


def test_both_debye_hueckel_variants_share_the_list(small_open):
    """D31: with both Debye-Hueckel variants enabled (equal screening
    distances, so the list is the same as with either alone) each term's
    energy and force equal its run alone: both evaluators run over the one
    list."""
    r = _compressed(small_open, 0.8)
    res = {}
    for name, terms in (("const", {"const_es": True}), ("rel", {"rel_es": True}),
                        ("both", {"const_es": True, "rel_es": True})):
        p = _single_term(1, "tether")
        synth.apply_terms(p, dict(tether=False, **terms))
        ora = Oracle(p, small_open)
        ora.set_positions(0, r)
        ora.rebuild_lists()
        res[name] = {t: ora.eval_forces(1 << _abi.TERM[t]) for t in ("const_es", "rel_es")}
    kc, kr = _abi.TERM["const_es"], _abi.TERM["rel_es"]
    assert res["both"]["const_es"][1][0, kc] == res["const"]["const_es"][1][0, kc] != 0
    assert res["both"]["rel_es"][1][0, kr] == res["rel"]["rel_es"][1][0, kr] != 0
    assert np.array_equal(res["both"]["rel_es"][0], res["rel"]["rel_es"][0])
    assert np.array_equal(res["both"]["const_es"][0], res["const"]["const_es"][0])
