"""Observables of a replica's state (host side, numpy).

  Rg: xmd/src/stats/gyration.cpp:4-18 (unit masses)
  W:  xmd/src/stats/asphericity.cpp (singular values of the centred
      3 x n positions, W = 1.5 s_min^2 - 0.5 |s|^2)
  K:  xmd/src/stats/total_energy.cpp:4-10
  Q:  fraction of native contacts active, |r1 - r2| < thr 2^(-1/6) d_nat
      with thr = "active threshold" (1.5),
      xmd/src/io/report_structure.cpp:81-85
"""
import numpy as np


def gyration_radius(r):
    """r: (3, n) positions."""
    com = r.mean(axis=1, keepdims=True)
    return float(np.sqrt(((r - com) ** 2).sum() / r.shape[1]))


def asphericity(r):
    """r: (3, n) positions."""
    s = np.linalg.svd(r - r.mean(axis=1, keepdims=True), compute_uv=False)
    return float(1.5 * s[-1] ** 2 - 0.5 * (s ** 2).sum())


def kinetic_energy(v):
    return float(0.5 * (v ** 2).sum())


def native_fraction(r, nat_i1, nat_i2, nat_dist, active_threshold=1.5):
    if len(nat_i1) == 0:
        return 1.0
    d = np.linalg.norm(r[:, nat_i1] - r[:, nat_i2], axis=0)
    return float(np.mean(d < active_threshold * 2.0 ** (-1.0 / 6.0) * nat_dist))


def replica_observables(host, system, rep):
    """Rg, K, Q of replica `rep` of an engine or oracle driver."""
    y, v, _, _ = host.get_state(rep)
    sl = system.replica_slice(rep)
    m = (system.nat_i1 >= sl.start) & (system.nat_i1 < sl.stop)
    return (gyration_radius(y[0]), kinetic_energy(v),
            native_fraction(y[0], system.nat_i1[m] - sl.start, system.nat_i2[m] - sl.start, system.nat_dist[m]))
