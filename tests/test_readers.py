"""spmReadDriver on the reference's own test matrices (test/matrix/*, used by
test/CMakeLists.txt:144-153 and example/CTestLists.txt:117-175), kept as
fixtures under tests/golden/matrix/.

Each reader is checked against an independent reading of the same file:
scipy.io.mmread for the Matrix Market files, and a Python restatement of the
Harwell-Boeing fixed-width layout for the HB / RSA files (scipy's reader
refuses both: one is symmetric, the other has a right-hand-side format
field)."""
import os
import re

import numpy as np
import pytest
import scipy.io as sio
import scipy.sparse as sp

import pastix_amd as px

MAT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "matrix")


def _full(spm):
    cp, rp, v = spm.arrays()
    n = spm.n
    A = sp.csc_matrix((v, rp, cp), shape=(n, n))
    if spm.s.mtxtype == px.SpmSymmetric:
        A = A + A.T - sp.diags(A.diagonal())
    elif spm.s.mtxtype == px.SpmHermitian:
        A = A + A.conj().T - sp.diags(A.diagonal())
    return A.tocsc()


def _hb_python(path):
    """Harwell-Boeing by fixed-width fields (Duff, Grimes, Lewis 1992)."""
    with open(path) as f:
        lines = f.read().split("\n")
    cards = [int(t) for t in lines[1].split()]
    rhscrd = cards[4] if len(cards) > 4 else 0
    mxtype = lines[2][:3].upper()
    nrow, ncol, nnz = [int(t) for t in lines[2][3:].split()[:3]]
    fmts = re.findall(r"\(([^)]*)\)", lines[3])
    start = 5 if rhscrd > 0 else 4

    def fmt(s):
        m = re.search(r"(\d*)[IEDFG](\d+)", s.upper())
        return int(m.group(1) or 1), int(m.group(2))

    def take(count, pos, f, conv):
        per, w = fmt(f)
        out = []
        while len(out) < count:
            ln = lines[pos]
            pos += 1
            for c in range(per):
                fld = ln[c * w:(c + 1) * w]
                if not fld.strip() or len(out) >= count:
                    break
                out.append(conv(fld))
        return out, pos

    ptr, pos = take(ncol + 1, start, fmts[0], int)
    ind, pos = take(nnz, pos, fmts[1], int)
    val, pos = take(nnz, pos, fmts[2], lambda t: float(t.upper().replace("D", "E")))
    A = sp.csc_matrix((np.array(val), np.array(ind) - 1, np.array(ptr) - 1), shape=(nrow, ncol))
    if mxtype[1] == "S":
        A = A + A.T - sp.diags(A.diagonal())
    return A.tocsc(), mxtype


@pytest.mark.parametrize("name", ["young4c.mtx", "mhd1280b.mtx"])
def test_matrix_market_reader_matches_scipy(name):
    spm = px.Spm.read(px.SpmDriverMM, os.path.join(MAT, name))
    ref = sio.mmread(os.path.join(MAT, name)).tocsc()
    A = _full(spm)
    assert spm.is_complex
    assert spm.s.mtxtype == (px.SpmSymmetric if name.startswith("young") else px.SpmHermitian)
    assert A.shape == ref.shape
    assert abs(A - ref).max() == 0
    # rows sorted within each column, 0-based
    cp, rp, v = spm.arrays()
    assert all(np.all(np.diff(rp[cp[j]:cp[j + 1]]) > 0) for j in range(spm.n))


@pytest.mark.parametrize("name,driver", [("orsirr.rua", "SpmDriverHB"), ("small.rsa", "SpmDriverRSA"),
                                         ("small.rsa", "SpmDriverHB")])
def test_harwell_boeing_reader_matches_fixed_width_reading(name, driver):
    spm = px.Spm.read(getattr(px, driver), os.path.join(MAT, name))
    ref, mxtype = _hb_python(os.path.join(MAT, name))
    assert spm.s.mtxtype == (px.SpmSymmetric if mxtype[1] == "S" else px.SpmGeneral)
    assert not spm.is_complex
    A = _full(spm)
    assert abs(A - ref).max() == 0


def test_laplacian_driver_string():
    spm = px.Spm.read(px.SpmDriverLaplacian, "d:6:5:4")
    ref = px.Spm.laplacian7(6, 5, 4, 6.0, 1.0)
    assert abs(_full(spm) - _full(ref)).max() == 0
    spm2 = px.Spm.read(px.SpmDriverLaplacian, "7:7:1:2.:0.5")      # 2-D grid: diagonal 2 * 4
    A = _full(spm2).toarray()
    assert spm2.n == 49 and A[0, 0] == 8.0 and A[1, 0] == -0.5
    with pytest.raises(px.PastixError):
        px.Spm.read(px.SpmDriverLaplacian, "10:x")


def test_ijv_reader(tmp_path):
    p = tmp_path / "m.ijv"
    p.write_text("3 4\n1 1 2.0\n2 2 3.0\n3 1 -1.0\n3 3 4.5\n")
    spm = px.Spm.read(px.SpmDriverIJV, p)
    A = _full(spm).toarray()
    assert np.array_equal(A, [[2, 0, 0], [0, 3, 0], [-1, 0, 4.5]])


def test_reader_errors(tmp_path):
    with pytest.raises(px.PastixError) as e:
        px.Spm.read(px.SpmDriverMM, tmp_path / "missing.mtx")
    assert e.value.code == px.PASTIX_ERR_FILE
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1.0\n")
    with pytest.raises(px.PastixError) as e:
        px.Spm.read(px.SpmDriverMM, bad)
    assert e.value.code == px.PASTIX_ERR_FILE
    with pytest.raises(px.PastixError) as e:
        px.Spm.read(px.SpmDriverGraph, bad)
    assert e.value.code == px.PASTIX_ERR_NOTIMPLEMENTED
