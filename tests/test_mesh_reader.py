"""CPU tests of the T-FlowS mesh reader (csrc/meshio.cpp through
krylov.Mesh): NAME.cns (Process/CnsLoa.f90:38-62), NAME.geo
(Library/GeoLoa.f90:22-49), the pattern of Process/TopolM.f90, the centre
distances of Process/Calc3.f90 and TypeBC of Process/BouLoa.f90.

The fixtures under tests/golden/meshes were written by the reference's own
mesh generator (tests/golden/make_mesh_fixtures.py), whose printed grid
statistics (cells, sides, boundary cells, matrix entries, neighbour
histogram) pin the sizes and the pattern independently of the reader.  The
geometry is pinned by what the mesh must satisfy: closed cells, the
divergence-theorem volume, boundary cells sitting on their face centres,
periodic shifts that bring the partner cell next to its face.
"""
import json
import os
import struct

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")
CASES = ["1solid", "cavity_small", "channel_periodic", "2mat_small"]
# boundary types typBou(1:NB) of the .b files (Generate/Library_cases/1solid.b,
# cavity.b, 2mat.b; the gen.b of Test_cases/Problem1/problem1.sh), codes of
# allp_mod; StateMat of 2mat.b (FLUID 7, SOLID 8)
TYPBOU = {"1solid": [4, 2, 2], "cavity_small": [1, 1, 4], "channel_periodic": [1, 3], "2mat_small": [2, 1, 3]}
STATEMAT = {"2mat_small": [7, 8]}


def _mesh(pkg, name, geo=True):
    base = os.path.join(MESHES, name)
    return pkg.krylov.Mesh(base + ".cns", base + ".geo" if geo else None)


def _stats(name):
    return json.load(open(os.path.join(MESHES, name + ".json")))


@pytest.mark.parametrize("name", CASES)
def test_sizes_and_pattern_match_generator(pkg, name):
    m = _mesh(pkg, name)
    st = _stats(name)
    assert (m.nc, m.ns, m.nbc) == (st["nc"], st["ns"], st["nbc"])
    assert m.nonz == st["nnz"] + 1                                     # TopolM.f90:27
    rows = np.diff(m.acol)
    hist = {int(k): int(v) for k, v in zip(*np.unique(rows - 1, return_counts=True))}
    assert hist == {int(k): v for k, v in st["neighbours"].items()}
    assert m.acol[0] == 1 and m.acol[-1] == m.nonz and m.arow[-1] == 0
    # every row holds its diagonal at Adia and sorted columns
    for c in range(1, m.nc + 1):
        r = m.arow[m.acol[c - 1] - 1:m.acol[c] - 1]
        assert np.all(np.diff(r) >= 0) and m.arow[m.adia[c - 1] - 1] == c
    inner = m.sidec[2] > 0
    if st["periodic_links"]:
        # periodic faces join cells across the domain: their file D is the shift
        shift = np.abs(m.dx_file) + np.abs(m.dy_file) + np.abs(m.dz_file)
        assert int(np.count_nonzero(shift[inner])) == st["periodic_links"]
    m.close()


@pytest.mark.parametrize("name", CASES)
def test_pattern_equals_topolm_restatement(pkg, name):
    """The C++ TopolM against meshgen.topolm (the Python restatement used by
    the system generators) on the interior faces."""
    m = _mesh(pkg, name, geo=False)
    inner = m.sidec[2] > 0
    t = pkg.meshgen.topolm(m.nc, m.sidec[1][inner], m.sidec[2][inner])
    assert np.array_equal(t.acol, m.acol) and np.array_equal(t.arow, m.arow) and np.array_equal(t.adia, m.adia)
    assert np.array_equal(t.sidaij + 1, m.sidaij[:, inner])
    assert np.all(m.sidaij[:, ~inner] == 0)
    m.close()


@pytest.mark.parametrize("name", CASES)
def test_geometry_is_consistent(pkg, name):
    m = _mesh(pkg, name)
    nb, nc = m.nbc, m.nc
    c1, c2 = m.sidec[1].astype(int), m.sidec[2].astype(int)
    inner = c2 > 0
    S = np.stack([m.sx, m.sy, m.sz], 1)
    area = np.linalg.norm(S, axis=1)
    # closed cells: the outward areas of every cell sum to zero
    tot = np.zeros((nc, 3))
    np.add.at(tot, c1 - 1, S)
    np.add.at(tot, c2[inner] - 1, -S[inner])
    assert np.abs(tot).max() <= 1e-12 * area.max()
    # the volume by the divergence theorem, the face centre seen from c2
    # shifted back by the periodic offset
    xf = np.stack([m.xsp, m.ysp, m.zsp], 1)
    sh = np.stack([m.dx_file, m.dy_file, m.dz_file], 1)
    vol = np.zeros(nc)
    np.add.at(vol, c1 - 1, np.einsum("ij,ij->i", xf, S) / 3.0)
    np.add.at(vol, c2[inner] - 1, -np.einsum("ij,ij->i", xf[inner] - sh[inner], S[inner]) / 3.0)
    V = m.volume[nb + 1:]
    assert np.all(V > 0)
    np.testing.assert_allclose(vol, V, rtol=1e-9, atol=1e-12 * V.max())
    # boundary cells sit on their faces; cell arrays in (-NbC:NC) order
    xc = np.stack([m.xc, m.yc, m.zc], 1)
    assert np.allclose(xc[nb + c2[~inner]], xf[~inner], rtol=0, atol=1e-12 * np.abs(xf).max())
    # Calc3's centre distance and Scoef
    D = np.stack([m.dx, m.dy, m.dz], 1)
    np.testing.assert_array_equal(D, xc[nb + c2] + sh - xc[nb + c1])
    sc = ((m.sx * m.sx + m.sy * m.sy) + m.sz * m.sz) / ((m.dx * m.sx + m.dy * m.sy) + m.dz * m.sz)
    np.testing.assert_array_equal(m.scoef, sc)
    assert np.all(m.scoef > 0)
    assert np.all((m.f >= 0) & (m.f <= 1))
    # the periodic partner lands next to c1: |D| is a cell size, not the domain
    assert np.linalg.norm(D, axis=1).max() < 0.5 * (xc[nb + 1:].max(0) - xc[nb + 1:].min(0)).max() + 1e-12
    m.close()


@pytest.mark.parametrize("name", CASES)
def test_materials_marks_typebc_ff(pkg, name):
    m = _mesh(pkg, name)
    nb = m.nbc
    assert m.nmat >= 1 and np.all((m.material[nb + 1:] >= 1) & (m.material[nb + 1:] <= m.nmat))
    assert np.all(m.bcmark[:nb] >= 1) and np.all(m.bcmark[nb:] == 0)
    typ = TYPBOU[name]
    tb = m.typebc(typ)
    want = np.zeros_like(tb)
    marks = m.bcmark[:nb]
    want[:nb] = np.where((marks >= 1) & (marks <= len(typ)), np.asarray(typ)[np.clip(marks, 1, len(typ)) - 1], 0)
    want[:nb][marks == 11] = 11
    assert np.array_equal(tb, want)
    st = STATEMAT.get(name)
    ff = m.ff(tb, st)
    bnd = m.sidec[2] < 0
    assert np.all(ff[bnd] == 1.0)
    mat = m.material
    m1, m2 = mat[nb + m.sidec[1]], mat[nb + m.sidec[2]]
    mixed = ~bnd & (m1 != m2)
    assert np.array_equal(ff[~bnd & ~mixed], m.f[~bnd & ~mixed])
    if st is None:
        assert np.array_equal(ff[mixed], m.f[mixed])
    else:
        # fluid c1 next to solid c2: 1; solid c1 next to fluid c2: 0
        assert mixed.any()
        s1, s2 = np.asarray(st)[m1[mixed] - 1], np.asarray(st)[m2[mixed] - 1]
        assert np.array_equal(ff[mixed], np.where((s1 == 7) & (s2 == 8), 1.0, 0.0))
    m.close()


def _record(payload: bytes) -> bytes:
    return struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload))


def _write_cns(path, nc, nbc, sidec, nmat=1, int8=False, split=None):
    fmt = "<q" if int8 else "<i"
    ints = lambda xs: b"".join(struct.pack(fmt, int(x)) for x in xs)   # noqa: E731
    recs = [ints([nc]), ints([nbc]), ints([sidec.shape[1]]), ints([0]), ints([nmat]), ints([1] * nc),
            ints([1] * nbc), ints(sidec[0]), ints(sidec[1]), ints(sidec[2]), ints([1] * nbc), ints([0] * nbc)]
    out = b""
    for i, r in enumerate(recs):
        if split is not None and i == split and len(r) > 8:
            # one record as two subrecords: the first count negated
            a, b = r[:8], r[8:]
            out += struct.pack("<i", -len(a)) + a + struct.pack("<i", -len(a)) + _record(b)
        else:
            out += _record(r)
    open(path, "wb").write(out)


def test_duplicate_face_pairs_follow_topolm(pkg, tmp_path):
    """Two faces joining the same two cells: TopolM counts both in the row
    but its search gives both faces the last matching position (:65-74)."""
    sidec = np.array([[0, 0, 0, 0], [1, 1, 2, 1], [2, 2, 3, -1]])
    p = str(tmp_path / "dup.cns")
    _write_cns(p, 3, 1, sidec)
    m = pkg.krylov.Mesh(p)
    assert m.nonz == 3 + 2 * 3 + 1
    assert list(m.acol) == [1, 4, 8, 10]
    assert list(m.arow) == [1, 2, 2, 1, 1, 2, 3, 2, 3, 0]
    assert list(m.sidaij[0]) == [3, 3, 7, 0] and list(m.sidaij[1]) == [5, 5, 8, 0]
    m.close()


def test_record_kinds_and_subrecords(pkg, tmp_path):
    """8-byte integers (another build's kind) and a record split into
    subrecords read the same as the plain file."""
    sidec = np.array([[0, 0, 0], [1, 2, 1], [2, 3, -1]])
    ref = None
    for i, kw in enumerate([dict(), dict(int8=True), dict(split=7)]):
        p = str(tmp_path / f"k{i}.cns")
        _write_cns(p, 3, 1, sidec, **kw)
        m = pkg.krylov.Mesh(p)
        got = (m.nc, m.nbc, m.ns, list(m.sidec.ravel()), list(m.arow))
        ref = ref or got
        assert got == ref
        m.close()


def test_bad_files_raise(pkg, tmp_path):
    good = os.path.join(MESHES, "channel_periodic.cns")
    data = open(good, "rb").read()
    cut = tmp_path / "cut.cns"
    cut.write_bytes(data[:len(data) // 2])
    with pytest.raises(pkg.krylov.TFlowsError, match="record|end of file"):
        pkg.krylov.Mesh(str(cut))
    # a header that claims 2^31 - 1 bytes in a short file is refused before
    # any buffer of that size is made
    huge = tmp_path / "huge.cns"
    huge.write_bytes(struct.pack("<i", 2**31 - 1) + b"\x00" * 64)
    with pytest.raises(pkg.krylov.TFlowsError, match="shorter than its count"):
        pkg.krylov.Mesh(str(huge))
    bad = tmp_path / "bad.cns"
    bad.write_bytes(b"\x07\x00\x00\x00not a mesh\x00")
    with pytest.raises(pkg.krylov.TFlowsError):
        pkg.krylov.Mesh(str(bad))
    with pytest.raises(pkg.krylov.TFlowsError, match="cannot open"):
        pkg.krylov.Mesh(str(tmp_path / "missing.cns"))
    # a .geo that does not belong to the .cns
    with pytest.raises(pkg.krylov.TFlowsError):
        pkg.krylov.Mesh(good, os.path.join(MESHES, "1solid.geo"))
