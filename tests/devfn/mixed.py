"""Per-term FP32-vs-FP64 force error of the engine's physics functions on a
given conformation, evaluated on the CPU through tests/devfn/libdevfn.so
(a host build of xmd_amd/csrc/xmd_dev.hpp).  Displacements are formed in
FP64 and narrowed, as the mixed kernels do; per-bead sums are FP64."""
import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_L = None


def lib():
    global _L
    if _L is None:
        subprocess.run(["make", "-s", "-C", HERE], check=True)
        _L = C.CDLL(os.path.join(HERE, "libdevfn.so"))
    return _L


def _p(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))


def params_vec(p):
    es_rel = 0 if p.const_es_enabled else 1
    Vf = 1 / (4 * math.pi * (p.const_es_permittivity if not es_rel else p.rel_es_factor))
    sdi = 1 / (p.const_es_screen if not es_rel else p.rel_es_screen)
    return np.array([p.tether_H1, p.tether_H2, p.chir_e, p.nat_ang_k, p.cnd_CDA, p.cnd_CDB, p.snd_CDH,
                     p.pauli_r_excl, p.pauli_depth, p.go_depth, Vf, sdi, es_rel,
                     p.pid_bbp_psi0, p.pid_bbp_alpha, p.pid_bbp_rmin, p.pid_bbp_depth,
                     p.pid_bbm_psi0, p.pid_bbm_alpha, p.pid_bbm_rmin, p.pid_bbm_depth,
                     p.pid_ss_psi0, p.pid_ss_alpha, p.pid_lambda_version], dtype=np.float64)


def wrap(d, L):
    Linv = np.where(L != 0, 1.0 / np.where(L != 0, L, 1), 0.0)
    return d - L * np.rint(d * Linv)


def term_forces(p, system, r, lists, prec):
    """dict term -> (n, 3) FP64-summed forces from FP32 (prec=1) or FP64
    (prec=0) per-item math.  r: (n, 3); lists: oracle lists by name."""
    L = lib()
    n = r.shape[0]
    box = system.box[0]
    prm = params_vec(p)
    out = {}

    def acc(idx, F):
        G = np.zeros((n, 3))
        for k in range(idx.shape[1]):
            np.add.at(G, idx[:, k], F[:, k])
        return G

    # PID
    i1, i2 = lists["pid"]
    if len(i1):
        a24 = wrap(r[i2] - r[i1], box)
        geo = np.stack([r[i1] - r[i1 - 1], r[i1 + 1] - r[i1], r[i1 + 1] - r[i1 - 1], a24,
                        r[i2] - r[i2 - 1], r[i2 + 1] - r[i2], r[i2 + 1] - r[i2 - 1], -a24, a24], axis=1)
        ty = 20 * system.atype[i1].astype(int) + system.atype[i2].astype(int)
        ss_rmax = np.array(p.ss_rmax)[ty]
        ss = np.stack([np.array(p.ss_depth)[ty], np.full(len(ty), p.lj_bb_rmin) if p.ss_sinking else ss_rmax,
                       ss_rmax], axis=1)
        o = np.zeros((len(i1), 19))
        L.h_pid(prec, len(i1), _p(geo.reshape(-1)), _p(ss.reshape(-1)), _p(prm), o.ctypes.data_as(C.POINTER(C.c_double)))
        idx = np.stack([i1 - 1, i1, i1 + 1, i2 - 1, i2, i2 + 1], axis=1)
        out["pid"] = acc(idx, o[:, :18].reshape(-1, 6, 3))
    # bonded quads
    chain = np.searchsorted(system.chain_offset, np.arange(n), side="right") - 1
    q = np.flatnonzero((np.arange(n) + 3 < n) & (chain == np.roll(chain, -3)) & np.isfinite(system.nat_phi))
    geo = np.stack([r[q + 1] - r[q], r[q + 2] - r[q + 1], r[q + 3] - r[q + 2]], axis=1)
    idx = np.stack([q, q + 1, q + 2, q + 3], axis=1)
    for kind, name, nat in ((0, "cnd", system.nat_phi[q]),
                            (2, "chiral", np.stack([system.nat_chir[q], system.nat_chir_factor[q]], 1))):
        o = np.zeros((len(q), 13))
        L.h_quad(prec, kind, len(q), _p(geo.reshape(-1)), _p(nat.reshape(-1)), _p(prm),
                 o.ctypes.data_as(C.POINTER(C.c_double)))
        out[name] = acc(idx, o[:, :12].reshape(-1, 4, 3))
    a = np.flatnonzero((np.arange(n) + 2 < n) & (chain == np.roll(chain, -2)) & np.isfinite(system.nat_theta))
    geo = np.stack([r[a + 1] - r[a], r[a + 2] - r[a + 1]], axis=1)
    o = np.zeros((len(a), 10))
    L.h_angle(prec, len(a), _p(geo.reshape(-1)), _p(system.nat_theta[a]), _p(prm), o.ctypes.data_as(C.POINTER(C.c_double)))
    out["nat_ang"] = acc(np.stack([a, a + 1, a + 2], 1), o[:, :9].reshape(-1, 3, 3))
    # pairs
    t = np.flatnonzero((np.arange(n) + 1 < n) & (chain == np.roll(chain, -1)))
    pairs = {"tether": (0, t, t + 1, np.full(len(t), p.tether_def_length), False)}
    g1, g2 = lists["go"]
    nat = {(a, b): d for a, b, d in zip(system.nat_i1, system.nat_i2, system.nat_dist)}
    pairs["go"] = (1, g1, g2, np.array([nat[(x, y)] for x, y in zip(g1, g2)]), True)
    pa1, pa2 = lists["pauli"]
    pairs["pauli"] = (2, pa1, pa2, np.zeros(len(pa1)), True)
    e1, e2 = lists["es"]
    qv = np.array(p.aa_charge)
    pairs["const_es"] = (3, e1, e2, qv[system.atype[e1]] * qv[system.atype[e2]], True)
    for name, (kind, x, y, par, pbc) in pairs.items():
        if len(x) == 0:
            continue
        d = r[y] - r[x]
        if pbc:
            d = wrap(d, box)
        o = np.zeros((len(x), 4))
        L.h_pair(prec, kind, len(x), _p(d.reshape(-1)), _p(par), _p(prm), o.ctypes.data_as(C.POINTER(C.c_double)))
        f = o[:, :3]
        out[name] = acc(np.stack([x, y], 1), np.stack([f, -f], 1))
    return out
