"""YAML front end: units, merge semantics and the parameter image."""
import math

import pytest

from xmd_amd import params as P, _abi


def test_unit_expressions():
    A = 0.2
    assert P.quantity("3.8 A") == pytest.approx(3.8 * A)
    assert P.quantity("100 eps/A**2") == pytest.approx(100 / (A * A))
    assert P.quantity("1e-3 atom/A**3") == pytest.approx(1e-3 / A ** 3)
    assert P.quantity("2.0 1/tau") == 2.0
    assert P.quantity("0.38 eps/kB") == 0.38
    assert P.quantity("30.0 eps/rad**2") == 30.0
    assert P.quantity("180 deg") == pytest.approx(math.pi)
    assert P.quantity("4.5", A) == pytest.approx(0.9)
    assert P.quantity("(1+2)*3 A") == pytest.approx(9 * A)


def test_permittivity_gives_reference_debye_prefactor():
    """eps_0 from SI constants in the internal units: 1/(4 pi 80 eps_0) is
    332.06 kcal/mol*A/80 = 0.526 eps*f77unit for unit charges."""
    e0 = P.quantity("80.0 eps_0")
    assert 1 / (4 * math.pi * e0) == pytest.approx(0.526, rel=1e-3)
    kcal = P.UNITS["kcal"] / P.UNITS["mol"]  # per molecule
    coulomb_const = 332.0637 * kcal * P.UNITS["A"]  # kcal/mol * A in internal units
    assert 1 / (4 * math.pi * P.UNITS["eps_0"]) == pytest.approx(coulomb_const, rel=1e-4)


def test_merge_yaml_overrides_leaves():
    base = {"a": {"b": 1, "c": 2}, "d": 3}
    out = P.merge_yaml(base, {"a": {"c": 5}, "e": 6})
    assert out == {"a": {"b": 1, "c": 5}, "d": 3, "e": 6}
    assert base["a"]["c"] == 2


def test_defaults_to_params():
    p, tree = P.from_yaml()
    A = 0.2
    assert p.dt == pytest.approx(5e-3)
    assert p.gamma_factor == 2.0
    assert p.pad_factor == 0.25
    assert p.tether_H1 == pytest.approx(100 / A ** 2)
    assert p.tether_def_length == pytest.approx(3.8 * A)
    assert p.pauli_r_excl == pytest.approx(4 * A)
    assert p.lj_bb_rmin == pytest.approx(5 * A) and p.lj_bs_rmin == pytest.approx(6.8 * A)
    k = lambda a, b: 20 * P.AA_INDEX[a] + P.AA_INDEX[b]
    assert p.ss_rmax[k("LEU", "LEU")] == pytest.approx(8.93237 * A)
    assert p.ss_rmax[k("TRP", "TRP")] == pytest.approx(10.8456 * A)
    assert p.ss_rmax[k("GLY", "ALA")] == 0.0
    # symmetric table
    for a in P.AA_NAMES:
        for b in P.AA_NAMES:
            assert p.ss_rmax[k(a, b)] == p.ss_rmax[k(b, a)]
    assert p.aa_charge[P.AA_INDEX["LYS"]] == 1 and p.aa_charge[P.AA_INDEX["ASP"]] == -1
    assert list(p.aa_limits[P.AA_INDEX["TRP"]]) == [2, 8, 3, 4]
    assert p.aa_polarization[P.AA_INDEX["GLY"]] == 0 and p.aa_polarization[P.AA_INDEX["SER"]] == 1
    assert p.qa_enabled == 1 and p.pid_enabled == 0 and p.const_es_enabled == 1
    assert p.pid_lambda_version == _abi.LAMBDA_COSINE


def test_user_file_overrides(tmp_path):
    f = tmp_path / "in.yml"
    f.write_text("langevin:\n  temperature: 0.5 eps/kB\nintegrator:\n  dt: 2e-3 tau\n"
                 "pseudo-improper dihedral potential:\n  enabled: true\n  lambda version: algebraic\n")
    p, tree = P.from_yaml(str(f))
    assert p.dt == pytest.approx(2e-3)
    assert p.pid_enabled == 1 and p.pid_lambda_version == _abi.LAMBDA_ALGEBRAIC
    assert P.quantity(tree["langevin"]["temperature"]) == 0.5
