"""File reporters in the reference's formats (report_stats.cpp,
report_structure.cpp, export_pdb.cpp + pdb.cpp record writers): exact
strings on hand-built inputs, and PDB output read back by the loader."""
import math
import os

import numpy as np
import pytest

from xmd_amd import model as M, params as P, reporters as Rp

A = P.ANGSTROM


def small_model():
    m = M.Model()
    pos = np.array([[0.0, 0.0, 0.0], [3.8, 0.0, 0.0], [3.8, 3.8, 0.0], [7.6, 3.8, 1.0], [7.6, 0.0, 1.0]]) * A
    m.chains = [[m.add_residue(P.AA_INDEX[n], p) for n, p in zip(["MET", "GLY", "CYS"], pos[:3])],
                [m.add_residue(P.AA_INDEX[n], p) for n, p in zip(["ALA", "TRP"], pos[3:])]]
    m.contacts = [(0, 3, 8.0 * A, M.UNKNOWN), (1, 4, 5.0 * A, M.SIDE_BACK)]
    m.angles = [(0, 1, 2, 1.5), (2, 3, 4, None)]
    m.dihedrals = [(0, 1, 2, 3, None)]
    return m, pos.T.copy()


def test_stats_csv(tmp_path):
    path = str(tmp_path / "out.csv")
    rep = Rp.StatsReport(path, period=10.0)
    # t, V, K, E, Rg, W (internal units), active, total
    assert rep([0.0, -1.5, 2.25, 0.75, 2.0, 0.5, 3, 4])
    assert not rep([5.0, 0, 0, 0, 1, 1, 0, 0])  # before the period
    assert rep([10.0, 1.0, 1.0, 2.0, 1.0, -0.25, 0, 0])
    lines = open(path).read().splitlines()
    assert lines[0] == "t[tau],V[eps],K[eps],E[eps],Rg[A],W[A]"
    # Rg 2 internal = 10 A; W / angstrom as the reference prints it
    assert lines[1] == "0.000000,-1.500000,2.250000,0.750000,10.000000,2.500000"
    assert lines[2] == "10.000000,1.000000,1.000000,2.000000,5.000000,-1.250000"
    assert len(lines) == 3


def test_structure_yaml(tmp_path):
    m, r = small_model()
    rep = Rp.StructureReport(str(tmp_path / "structure" / "%.3f.yml"), period=50.0)
    path = rep(12.5, m, r)
    assert path.endswith("structure/12.500.yml")
    text = open(path).read()
    lines = text.splitlines()
    assert lines[0] == "residues: |-"
    assert lines[1] == "  idx,x[A],y[A],z[A]"
    assert lines[3] == "  1,3.800000,0.000000,0.000000"
    i = lines.index("native contacts: |-")
    assert lines[i + 1] == "  idx,i1,i2,type,nat_dist[A],cur_dist[A],active"
    d03 = math.sqrt(7.6 ** 2 + 3.8 ** 2 + 1.0)  # 8.556 A < 1.5 2^(-1/6) 8 A = 10.69 A
    assert lines[i + 2] == "  0,0,3,,8.000000,%f,1" % d03
    d14 = math.sqrt(3.8 ** 2 + 1.0)  # 3.929 A < 6.68 A
    assert lines[i + 3] == "  1,1,4,SIDE_BACK,5.000000,%f,1" % d14
    j = lines.index("angles: |-")
    assert lines[j + 2] == "  0,0,1,2,1.500000,%f" % (math.pi / 2)
    assert lines[j + 3].startswith("  1,2,3,4,,")
    k = lines.index("dihedrals: |-")
    assert lines[k + 1] == "  idx,i1,i2,i3,i4,nat_phi[rad],cur_phi[rad]"
    assert lines[k + 2].startswith("  0,0,1,2,3,,")
    yaml_doc = __import__("yaml").safe_load(text)
    assert set(yaml_doc) == {"residues", "native contacts", "angles", "dihedrals"}


def test_pdb_export_round_trip(tmp_path):
    m, r = small_model()
    path = str(tmp_path / "out.pdb")
    rep = Rp.PdbExport(path, period=50.0)
    assert rep(0.0, m, r)
    assert rep(50.0, m, r + 0.1)
    text = open(path).read()
    assert not text.endswith("\n")
    lines = text.split("\n")
    assert lines[0].rstrip() == "MODEL        1"
    assert all(len(l) == 80 for l in lines)
    assert lines[1][:30] == "ATOM      1   CA MET A   1    "
    assert lines[1][30:54] == "   0.000   0.000   0.000"
    assert lines[4][:26] == "TER       0      CYS A   3"
    assert [l[:6] for l in lines].count("ENDMDL") == 2
    assert lines[lines.index("MODEL        2".ljust(80))].startswith("MODEL")
    # the loader reads the first model back (positions to the printed 3 decimals)
    first = "\n".join(lines[:lines.index("ENDMDL".ljust(80)) + 1])
    back = M.read_pdb(first + "\nEND   \n")
    assert [len(c) for c in back.chains] == [3, 2]
    assert [P.AA_NAMES[a] for a in back.atype] == ["MET", "GLY", "CYS", "ALA", "TRP"]
    assert np.abs(np.array(back.pos) - r.T).max() < 0.0006 * A


def test_structure_yaml_qa_sections():
    """The reporter's optional budget columns and QA contact section
    (report_structure.cpp:28-51,92-141) from get_qa_state's records."""
    from xmd_amd._host import CONTACT_DTYPE
    m, r = small_model()
    qa = np.zeros(2, dtype=CONTACT_DTYPE)
    qa[0] = (0, 3, 1, 0, 0, (1, 0, 0, 0), (1, 0, 0, 0), 3.5)
    qa[1] = (1, 4, 4 + 20 * 12 + 7, 1, 0, (0, 1, 0, 1), (0, 1, 1, 0), 12.25)
    sync = np.array([[2, 1, 0, 1], [1, 2, 2, 0], [0, 0, 0, 0], [1, 1, 1, 1], [3, 0, 0, 0]], dtype=np.int8)
    lines = Rp.structure_yaml(m, r, 1.5, qa_contacts=qa, sync=sync).splitlines()
    assert lines[1] == "  idx,x[A],y[A],z[A],back,side-all,side-polar,side-hydrophobic"
    assert lines[2] == "  0,0.000000,0.000000,0.000000,2,1,0,1"
    k = lines.index("QA contacts: |-")
    assert lines[k + 1] == "  idx,i1,i2,type,t0[tau],status"
    assert lines[k + 2] == "  0,0,3,BACK_BACK,3.500000,FORMING_OR_FORMED"
    assert lines[k + 3] == "  1,1,4,SIDE_SIDE,12.250000,BREAKING"
    assert lines[k - 1] == "" and lines[k - 2].startswith("  1,1,4,SIDE_BACK")  # after the native contacts
    assert lines[k + 5] == "angles: |-"
    # without the optional parts the report is the one main.cpp writes
    plain = Rp.structure_yaml(m, r, 1.5)
    assert "QA contacts" not in plain and "back" not in plain.splitlines()[1]
