"""The mixed-precision form on the CPU: the engine's FP32 physics functions
(xmd_amd/csrc/xmd_dev.hpp, host build in tests/devfn) against their FP64
form, which follows the reference expressions.

* atan2_f (a fitted polynomial) against numpy's arctan2;
* the cosine lambda through the half angle against 0.5 cos s + 0.5
  (lambda.h:30-42) in FP64;
* every term's FP32 forces against FP64 on an evolved periodic box (4 x 250
  beads at the workload's density, oracle trajectory), total rel-RMS within
  the 1e-5 the mixed form is held to, and the FP64 sums against the oracle;
* the angle and dihedral terms on the displaced 100k-bead box, whose
  near-straight angles need the plane normals formed in FP64.
"""
import ctypes as C
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "devfn"))
import mixed  # noqa: E402

from xmd_amd import _abi, synth  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

F32 = C.POINTER(C.c_float)
F64 = C.POINTER(C.c_double)


def atan2_f(y, x):
    y = np.ascontiguousarray(y, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(y)
    mixed.lib().h_atan2f(len(y), y.ctypes.data_as(F32), x.ctypes.data_as(F32), out.ctypes.data_as(F32))
    return out


def lam(prec, version, psi0, alpha, psi):
    psi = np.ascontiguousarray(psi, dtype=np.float64)
    out = np.empty(2 * len(psi))
    mixed.lib().h_lam(prec, version, C.c_double(psi0), C.c_double(alpha), len(psi), psi.ctypes.data_as(F64),
                      out.ctypes.data_as(F64))
    return out[0::2], out[1::2]


def test_atan2_f_against_numpy():
    rng = np.random.default_rng(1)
    ang = rng.uniform(-math.pi, math.pi, 200000)
    rad = np.exp(rng.uniform(-8, 8, ang.size))
    y = (rad * np.sin(ang)).astype(np.float32)
    x = (rad * np.cos(ang)).astype(np.float32)
    # octant borders, axes, signed zeros
    sy = np.array([0, -0.0, 1, -1, 1, -1, 0, -0.0, 1, 1, -1, 3, -3], dtype=np.float32)
    sx = np.array([1, 1, 0, 0, 1, 1, -1, -1, -1, -1e-30, -1, -3, 3], dtype=np.float32)
    y, x = np.concatenate([y, sy]), np.concatenate([x, sx])
    got = atan2_f(y, x).astype(np.float64)
    want = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert np.abs(got - want).max() < 4e-7
    assert np.signbit(got[-13 + 1]) and got[-13 + 7] == np.float32(-math.pi)  # atan2(-0, 1), atan2(-0, -1)


@pytest.mark.parametrize("psi0,alpha", [(1.05, 6.4), (-1.44, 6.0), (-0.23, 1.2)])
def test_cosine_lambda_half_angle(psi0, alpha):
    # the support |alpha (psi - psi0)| < pi
    psi = psi0 + np.linspace(-0.999, 0.999, 20001) * math.pi / alpha
    v32, d32 = lam(1, 0, psi0, alpha, psi.astype(np.float32).astype(np.float64))
    v64, d64 = lam(0, 0, psi0, alpha, psi.astype(np.float32).astype(np.float64))
    s = alpha * (psi.astype(np.float32).astype(np.float64) - psi0)
    assert np.abs(v64 - (0.5 * np.cos(s) + 0.5)).max() < 1e-15
    assert np.abs(v32 - v64).max() < 4e-7 * max(1.0, alpha)
    assert np.abs(d32 - d64).max() < 4e-7 * alpha * max(1.0, alpha)


def test_mixed_force_error_on_evolved_box():
    system = synth.config_system(2, chains=4)
    p = synth.config_params(2)
    ora = Oracle(p, system)
    ora.step(150)
    y, _, _, _ = ora.get_state(0)
    r = y[0].T.copy()
    lists = {k: ora.get_list(0, k) for k in ("pid", "go", "pauli", "es")}
    assert all(len(lists[k][0]) for k in ("pid", "go", "es"))
    Fd = mixed.term_forces(p, system, r, lists, 0)
    Ff = mixed.term_forces(p, system, r, lists, 1)
    tot_d, tot_f = sum(Fd.values()), sum(Ff.values())
    rel = math.sqrt(((tot_f - tot_d) ** 2).sum() / (tot_d ** 2).sum())
    assert rel < 1e-5, rel
    Fo, _ = ora.eval_forces(sum(1 << _abi.TERM[k] for k in Fd))
    assert math.sqrt(((tot_d - Fo.T) ** 2).sum() / (Fo ** 2).sum()) < 1e-12


def test_bonded_terms_near_straight_angles():
    """The 100k-bead box displaced by 0.3 A holds bond angles close to
    straight, where the plane normals of the angle and dihedral terms
    cancel; formed in FP64 and narrowed they keep each bonded term's FP32
    forces within 1e-6 rel-RMS of FP64 (FP32 normals: 8e-5 for the native
    dihedrals)."""
    system = synth.config_system(4)
    p = synth.config_params(4)
    rng = np.random.default_rng(3000)
    r = (system.r0 + rng.normal(scale=0.3 * synth.ANGSTROM, size=system.r0.shape)).T.copy()
    L, prm, n = mixed.lib(), mixed.params_vec(p), r.shape[0]
    chain = np.searchsorted(system.chain_offset, np.arange(n), side="right") - 1
    q = np.flatnonzero((np.arange(n) + 3 < n) & (chain == np.roll(chain, -3)) & np.isfinite(system.nat_phi))
    a = np.flatnonzero((np.arange(n) + 2 < n) & (chain == np.roll(chain, -2)) & np.isfinite(system.nat_theta))
    cases = [("cnd", q, 3, lambda prec, geo, o: L.h_quad(prec, 0, len(q), mixed._p(geo), mixed._p(system.nat_phi[q]),
                                                         mixed._p(prm), o.ctypes.data_as(F64))),
             ("snd", q, 3, lambda prec, geo, o: L.h_quad(prec, 1, len(q), mixed._p(geo), mixed._p(system.nat_phi[q]),
                                                         mixed._p(prm), o.ctypes.data_as(F64))),
             ("nat_ang", a, 2, lambda prec, geo, o: L.h_angle(prec, len(a), mixed._p(geo),
                                                              mixed._p(system.nat_theta[a]), mixed._p(prm),
                                                              o.ctypes.data_as(F64)))]
    for name, first, nb, call in cases:
        geo = np.stack([r[first + k + 1] - r[first + k] for k in range(nb)], axis=1).reshape(-1)
        F = {}
        for prec in (0, 1):
            o = np.zeros((len(first), 3 * (nb + 1) + 1))
            call(prec, geo, o)
            G = np.zeros((n, 3))
            for k in range(nb + 1):
                np.add.at(G, first + k, o[:, 3 * k:3 * k + 3])
            F[prec] = G
        rel = math.sqrt(((F[1] - F[0]) ** 2).sum() / (F[0] ** 2).sum())
        assert rel < 1e-6, (name, rel)
