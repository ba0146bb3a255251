"""CPU tests of the NAME.b reader (csrc/meshio.cpp through krylov.BFile,
after Process/BouLoa.f90:30-160 with the line rules of Library/ReadC.f90).

The fixtures in tests/golden/bfiles are boundary-condition files of the
reference's mesh-generator cases (Generate/Library_cases: 2mat.b, 1solid.b,
cond.b), copied as data.  Expected values are read off the files by hand;
2mat.b and 1solid.b belong to the mesh fixtures of test_mesh_reader, whose
boundary types and material states were typed there from the same files, so
reading them here must give the same TypeBC and fF.
"""
import os

import numpy as np
import pytest

from test_mesh_reader import MESHES, STATEMAT, TYPBOU

HERE = os.path.dirname(os.path.abspath(__file__))
BFILES = os.path.join(HERE, "golden", "bfiles")
INFLOW, WALL, OUTFLOW, SYMMETRY, CONVECT, WALLFL, PRESSURE = 1, 2, 3, 4, 5, 6, 12


def _b(pkg, name, hot=False):
    return pkg.krylov.BFile(os.path.join(BFILES, name), hot=hot)


def test_two_materials_file(pkg):
    b = _b(pkg, "2mat.b")
    assert (b.nmat, b.nb, b.ninit) == (2, 3, 2)
    assert b.statemat.tolist() == [7, 8] and b.denc.tolist() == [1.0, 1.0]
    assert b.visc == 0.005                      # one VISc, the last material's
    assert b.typbou.tolist() == [WALL, INFLOW, OUTFLOW]
    assert b.bfile.tolist() == [0, 1, 1]
    assert b.profiles == ["", "parabolic-0-2.prof", "parabolic-0-2.prof"]
    assert b.bu.tolist() == [0.0, 0.0, 0.0] and b.bp.tolist() == [0.0, 0.0, 0.0]
    assert b.ifile.tolist() == [0, 0] and b.inits == ["", ""]
    assert b.iu.tolist() == [1.0, 0.0] and b.iw.tolist() == [0.0, 0.0]
    assert np.all(b.iz == 0.0)                  # not HOT: the fifth number is not read


def test_mixed_case_types_and_hot_values(pkg):
    b = _b(pkg, "1solid.b")
    assert b.typbou.tolist() == [SYMMETRY, WALL, WALL] and b.statemat.tolist() == [8]
    assert b.visc == 0.05 and np.all(b.bz == 0.0)
    c = _b(pkg, "cond.b", hot=True)
    assert c.statemat.tolist() == [7] and c.conc.tolist() == [0.01] and c.capc.tolist() == [1.0]
    assert c.typbou.tolist() == [WALL, WALLFL, WALLFL, WALLFL]
    assert c.bz.tolist() == [0.0, 1.0, 0.0, 0.0]
    assert c.ninit == 1 and c.iz.tolist() == [0.0]
    # HOT asks for CONc and CAPc, which 2mat.b does not give
    with pytest.raises(pkg.krylov.TFlowsError, match="CONc"):
        _b(pkg, "2mat.b", hot=True)


@pytest.mark.parametrize("mesh,bname", [("2mat_small", "2mat.b"), ("1solid", "1solid.b")])
def test_types_feed_the_mesh(pkg, mesh, bname):
    """typBou and StateMat from the .b file give the mesh the TypeBC and fF
    of the types listed by hand for the same case."""
    b = _b(pkg, bname)
    assert b.typbou.tolist() == TYPBOU[mesh]
    m = pkg.krylov.Mesh(os.path.join(MESHES, mesh + ".cns"), os.path.join(MESHES, mesh + ".geo"))
    tb = m.typebc(b.typbou)
    assert np.array_equal(tb, m.typebc(TYPBOU[mesh]))
    if mesh in STATEMAT:
        assert b.statemat.tolist() == STATEMAT[mesh]
        assert np.array_equal(m.ff(tb, b.statemat), m.ff(tb, STATEMAT[mesh]))
    m.close()


def _write(tmp_path, text, name="t.b", newline="\n"):
    p = tmp_path / name
    p.write_bytes(text.replace("\n", newline).encode())
    return str(p)


GOOD = """% comment
! another
1
1 FLUID 1.0D-3 2.5
#
2
1 pressure 1.0 2.0 3.0 4.0
2 Convective FILE in.prof
1
1 FILE init.dat
"""


@pytest.mark.parametrize("newline", ["\n", "\r\n"])
def test_readc_rules(pkg, tmp_path, newline):
    """Comments by the first column, case-free keywords, a D exponent, the
    pressure value of a PRESSURE condition, FILE names; CRLF line ends."""
    b = pkg.krylov.BFile(_write(tmp_path, GOOD, newline=newline))
    assert b.visc == 1e-3 and b.denc.tolist() == [2.5]
    assert b.typbou.tolist() == [PRESSURE, CONVECT]
    assert (b.bu[0], b.bv[0], b.bw[0], b.bp[0]) == (1.0, 2.0, 3.0, 4.0)
    assert b.profiles == ["", "in.prof"] and b.inits == ["init.dat"] and b.ifile.tolist() == [1]


def test_readc_columns_and_blank_lines(pkg, tmp_path):
    """ReadC reads A300 (Library/ReadC.f90:21): a token past column 300 is
    never seen, so the density written there is missing; the same line
    within 300 columns reads normally.  Only a line of spaces is blank
    (:22): a tab-only line is a data line without tokens.  A token that
    reaches column 299 is never closed by ReadC's scan (:44-51) and is
    refused."""
    head = "1 FLUID 1.0D-3"
    inside = GOOD.replace("1 FLUID 1.0D-3 2.5", head + " " * (295 - len(head)) + "2.5")   # ends in column 298
    b = pkg.krylov.BFile(_write(tmp_path, inside, name="in.b"))
    assert b.visc == 1e-3 and b.denc.tolist() == [2.5]
    past = GOOD.replace("1 FLUID 1.0D-3 2.5", head + " " * (300 - len(head)) + "2.5")
    with pytest.raises(pkg.krylov.TFlowsError, match="missing"):
        pkg.krylov.BFile(_write(tmp_path, past, name="past.b"))
    edge = GOOD.replace("1 FLUID 1.0D-3 2.5", head + " " * (296 - len(head)) + "2.5")     # ends in column 299
    with pytest.raises(pkg.krylov.TFlowsError, match="column 299"):
        pkg.krylov.BFile(_write(tmp_path, edge, name="edge.b"))
    spaces = GOOD.replace("#\n", "    \n")
    assert pkg.krylov.BFile(_write(tmp_path, spaces, name="sp.b")).typbou.tolist() == [PRESSURE, CONVECT]
    tabs = GOOD.replace("#\n", "\t\n")
    with pytest.raises(pkg.krylov.TFlowsError, match="missing"):
        pkg.krylov.BFile(_write(tmp_path, tabs, name="tab.b"))


@pytest.mark.parametrize("text,match", [
    (GOOD.replace("FLUID", "GAS"), "material state"),
    (GOOD.replace("pressure", "WAL"), "boundary condition type"),
    (GOOD.replace("3.0 4.0", "3.0"), "missing P"),
    (GOOD.replace("2.0 3.0", "2.x 3.0"), "bad V"),
    (GOOD.replace("1\n1 FILE init.dat\n", ""), "end of file"),
    (GOOD.replace("#\n", " #\n"), "."),        # a comment mark after a blank is a token line
])
def test_bad_files(pkg, tmp_path, text, match):
    with pytest.raises(pkg.krylov.TFlowsError, match=match):
        pkg.krylov.BFile(_write(tmp_path, text))
    with pytest.raises(pkg.krylov.TFlowsError, match="cannot open"):
        pkg.krylov.BFile(str(tmp_path / "missing.b"))
