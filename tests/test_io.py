"""Ordering and symbol files, the command-line options, the dof expansion and
the built-in nested dissection, through the C API on the CPU.

- pastixOrderSave / pastixOrderLoad (reference order/order_io.c) through
  IPARM_IO_STRATEGY and directly: the file format is the reference's (a
  hand-written file in that format loads into the expected arrays), and a
  save followed by a load reproduces the ordering and the symbol structure;
- pastixSymbolSave / pastixSymbolLoad (symbol/symbol_io.c) round trip;
- pastixGetOptions (common/get_options.c) parses the reference's options;
- pastixExpand on a 3-dof problem gives the matrix of the scalar generator;
- the nested dissection that serves Scotch / Metis requests is a valid
  permutation with less fill than the natural order on grids and on the
  reference's test matrices.
"""
import ctypes
import glob
import os

import numpy as np
import pytest

import pastix_amd as px
from helpers import analyzed

MAT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "matrix")


def _order_struct(o):
    n, c = o.vertnbr, o.cblknbr
    return (np.ctypeslib.as_array(o.permtab, shape=(n,)).copy(), np.ctypeslib.as_array(o.rangtab, shape=(c + 1,)).copy(),
            np.ctypeslib.as_array(o.treetab, shape=(c,)).copy())


def test_order_file_format_is_the_reference_one(tmp_path, monkeypatch):
    # version 1, 3 supernodes over 5 vertices: rangtab, permtab, treetab
    (tmp_path / "ordername").write_text("1\n3\t5\n0\t2\t3\t5\n4\t3\t2\t1\t0\n1\t2\t-1\n")
    monkeypatch.chdir(tmp_path)
    monkeypatch.delenv("PASTIX_FILE_ORDER", raising=False)
    o = px.pastix_order_t()
    px._check("pastixOrderLoad", px.lib.pastixOrderLoad(None, ctypes.byref(o)))
    perm, rang, tree = _order_struct(o)
    assert list(perm) == [4, 3, 2, 1, 0] and list(rang) == [0, 2, 3, 5] and list(tree) == [1, 2, -1]
    peri = np.ctypeslib.as_array(o.peritab, shape=(5,))
    assert list(peri) == [4, 3, 2, 1, 0]
    px.lib.pastixOrderExit(ctypes.byref(o))
    # a version-0 file has no treetab
    (tmp_path / "v0").write_text("0\n1 3\n0 3\n2 0 1\n")
    monkeypatch.setenv("PASTIX_FILE_ORDER", str(tmp_path / "v0"))
    o = px.pastix_order_t()
    px._check("pastixOrderLoad", px.lib.pastixOrderLoad(None, ctypes.byref(o)))
    assert not bool(o.treetab)
    px.lib.pastixOrderExit(ctypes.byref(o))
    (tmp_path / "bad").write_text("2\n1 3\n")
    monkeypatch.setenv("PASTIX_FILE_ORDER", str(tmp_path / "bad"))
    o = px.pastix_order_t()
    assert px.lib.pastixOrderLoad(None, ctypes.byref(o)) == px.PASTIX_ERR_FILE


def _read_ints(path):
    return [int(t) for t in open(path).read().split()]


def test_order_and_symbol_saved_then_loaded_give_the_same_structure(tmp_path, monkeypatch):
    """IPARM_IO_STRATEGY PastixIOSave writes the final ordering ("ordergen",
    after the reordering) and symbol ("symbgen") into the output directory;
    PastixIOLoad reads them back ("ordername", "symbname"): the loaded
    ordering and symbol are the saved ones, and blend on them gives the same
    solver structure as the run that saved them."""
    nx = 7
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("PASTIX_OUTPUT_DIR", str(tmp_path / "out"))
    monkeypatch.delenv("PASTIX_FILE_ORDER", raising=False)
    spm = px.Spm.laplacian7(nx, nx, nx)
    pd, iparm, dparm = analyzed(spm, nx, extra={px.IPARM_IO_STRATEGY: px.PastixIOSave})
    s1 = px.solver_export(pd)
    o1 = px.order_export(pd)
    dirs = glob.glob(str(tmp_path / "out-*"))
    assert len(dirs) == 1
    assert os.path.exists(os.path.join(dirs[0], "ordergen")) and os.path.exists(os.path.join(dirs[0], "symbgen"))
    px.finalize(pd)
    # the ordering file holds version 1, cblknbr, vertnbr, rangtab, permtab, treetab
    v = _read_ints(os.path.join(dirs[0], "ordergen"))
    c, n = v[1], v[2]
    assert v[0] == 1 and n == spm.n
    assert np.array_equal(v[3:4 + c], o1[2]) and np.array_equal(v[4 + c:4 + c + n], o1[0])
    assert np.array_equal(v[4 + c + n:], o1[3])
    sy = _read_ints(os.path.join(dirs[0], "symbgen"))
    os.rename(os.path.join(dirs[0], "ordergen"), tmp_path / "ordername")
    os.rename(os.path.join(dirs[0], "symbgen"), tmp_path / "symbname")
    iparm, dparm = px.initParam()
    iparm[px.IPARM_ORDERING] = px.PastixOrderPersonal
    iparm[px.IPARM_IO_STRATEGY] = px.PastixIOLoad
    iparm[px.IPARM_VERBOSE] = px.PastixVerboseNot
    pd = px.init(iparm, dparm)
    px.subtask_order(pd, spm)
    o2 = px.order_export(pd)
    assert all(np.array_equal(a, b) for a, b in zip(o1, o2))
    px.subtask_symbfact(pd)
    cb, bk = px.symbol_export(pd)
    cblknbr, bloknbr = sy[1], sy[2]
    assert (cblknbr, bloknbr, sy[3]) == (len(cb) - 1, len(bk), spm.n)
    file_cb = np.array(sy[5:5 + 3 * cblknbr]).reshape(-1, 3)
    file_bk = np.array(sy[5 + 3 * cblknbr:]).reshape(-1, 3)
    assert np.array_equal(file_cb, cb[:-1]) and np.array_equal(file_bk, bk[:, [0, 1, 3]])
    px.subtask_blend(pd)
    s2 = px.solver_export(pd)
    assert all(np.array_equal(s1[k], s2[k]) for k in ("cblk", "blok", "browtab"))
    # pastixSymbolSave writes the same file again
    libc = ctypes.CDLL(None)
    libc.fopen.restype = ctypes.c_void_p
    libc.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    libc.fclose.argtypes = [ctypes.c_void_p]
    px.finalize(pd)
    pd = px.init(iparm, dparm)
    px.subtask_order(pd, spm)
    px.subtask_symbfact(pd)
    f = libc.fopen(str(tmp_path / "again").encode(), b"w")
    assert px.lib.pastixSymbolSave(px.lib.pastix_amd_symbol(pd), f) == 0
    libc.fclose(f)
    assert _read_ints(tmp_path / "again") == sy
    px.finalize(pd)
    # a missing symbol file is an error of the symbolic step
    os.remove(tmp_path / "symbname")
    pd = px.init(iparm, dparm)
    px.subtask_order(pd, spm)
    with pytest.raises(px.PastixError) as e:
        px.subtask_symbfact(pd)
    assert e.value.code == px.PASTIX_ERR_FILE
    px.finalize(pd)


def _argv(args):
    arr = (ctypes.c_char_p * (len(args) + 1))()
    for i, a in enumerate(args):
        arr[i] = a.encode()
    return arr


def test_get_options_parses_the_reference_options():
    iparm, dparm = px.initParam()
    check = ctypes.c_int(1)
    scatter = ctypes.c_int(0)
    driver = ctypes.c_int(-1)
    fname = ctypes.c_char_p()
    args = ["prog", "--mm", "young4c.mtx", "-f", "1", "-i", "iparm_refinement", "pastixrefinebicgstab",
            "-d", "dparm_epsilon_magn_ctrl", "1e-14", "-t", "3", "-g", "1", "-c", "2", "-s", "1",
            "-i", "IPARM_ITERMAX", "17"]
    px.lib.pastixGetOptions(len(args), _argv(args), iparm.ctypes.data_as(ctypes.c_void_p),
                            dparm.ctypes.data_as(ctypes.c_void_p), ctypes.byref(check), ctypes.byref(scatter),
                            ctypes.byref(driver), ctypes.byref(fname))
    assert driver.value == px.SpmDriverMM and fname.value == b"young4c.mtx"
    assert iparm[px.IPARM_FACTORIZATION] == 1 and iparm[px.IPARM_REFINEMENT] == px.PastixRefineBiCGSTAB
    assert dparm[px.DPARM_EPSILON_MAGN_CTRL] == 1e-14 and iparm[px.IPARM_ITERMAX] == 17
    assert iparm[px.IPARM_THREAD_NBR] == 3 and iparm[px.IPARM_GPU_NBR] == 1 and check.value == 2
    assert iparm[px.IPARM_SCHEDULER] == 1
    args = ["prog", "--rsa", "small.rsa"]
    px.lib.pastixGetOptions(len(args), _argv(args), iparm.ctypes.data_as(ctypes.c_void_p),
                            dparm.ctypes.data_as(ctypes.c_void_p), None, None, ctypes.byref(driver),
                            ctypes.byref(fname))
    assert driver.value == px.SpmDriverHB and fname.value == b"small.rsa"
    assert px.lib.pastix_amd_parse_iparm(b"iparm_gmres_im") == px.IPARM_GMRES_IM
    assert px.lib.pastix_amd_parse_dparm(b"DPARM_A_NORM") == px.DPARM_A_NORM
    assert px.lib.pastix_amd_parse_enum(b"pastixfactldlt") == px.PastixFactLDLT
    assert px.lib.pastix_amd_parse_enum(b"42") == 42 and px.lib.pastix_amd_parse_enum(b"nonsense") == -1


def test_expand_turns_a_dof_matrix_into_the_scalar_one():
    """A 3-dof node matrix (3x3 blocks: diagonal 6 c_i I, off-diagonal -c_ij I)
    expanded by pastixExpand equals the generator's scalar matrix."""
    nx, dof = 4, 3
    ref = px.Spm.vlaplacian7(nx, nx, nx, dof, 1.0, 0)
    cp, rp, v = ref.arrays()
    n = ref.n
    # node matrix: entry (I, J) of node pattern holds the block of rows/cols I*3.., J*3..
    import scipy.sparse as sp
    A = sp.csc_matrix((v, rp, cp), shape=(n, n)).toarray()
    nodes = n // dof
    ncp, nrp, nv = [0], [], []
    for J in range(nodes):
        for I in range(J, nodes):
            blk = A[I * dof:(I + 1) * dof, J * dof:(J + 1) * dof]
            if np.any(blk):
                nrp.append(I)
                nv.append(np.asfortranarray(blk).ravel(order="F"))
        ncp.append(len(nrp))
    m = px.Spm.from_csc(nodes, ncp, nrp, np.concatenate(nv), mtxtype=px.SpmSymmetric)
    # pastixExpand replaces the arrays (freeing the old ones): hand it malloc'ed copies
    libc = ctypes.CDLL(None)
    libc.malloc.restype = ctypes.c_void_p
    libc.malloc.argtypes = [ctypes.c_size_t]
    for field, arr in (("colptr", np.asarray(ncp, dtype=np.int64)), ("rowptr", np.asarray(nrp, dtype=np.int64)),
                       ("values", np.concatenate(nv))):
        ptr = libc.malloc(arr.nbytes)
        ctypes.memmove(ptr, arr.ctypes.data, arr.nbytes)
        setattr(m.s, field, ptr)
    m._keep = []
    m._owned_by_c = True
    m.s.dof = dof
    px.lib.pastixExpand(None, ctypes.byref(m.s))
    assert m.s.dof == 1 and m.n == n
    cp2 = np.ctypeslib.as_array(ctypes.cast(m.s.colptr, ctypes.POINTER(ctypes.c_int64)), shape=(n + 1,)).copy()
    nnz = cp2[-1]
    rp2 = np.ctypeslib.as_array(ctypes.cast(m.s.rowptr, ctypes.POINTER(ctypes.c_int64)), shape=(nnz,)).copy()
    v2 = np.ctypeslib.as_array(ctypes.cast(m.s.values, ctypes.POINTER(ctypes.c_double)), shape=(nnz,)).copy()
    B = sp.csc_matrix((v2, rp2, cp2), shape=(n, n)).toarray()
    B = np.tril(B)
    # the expanded matrix's lower triangle (explicit zeros of the blocks aside)
    assert np.array_equal(B, np.tril(A))


@pytest.mark.parametrize("name,driver", [("small.rsa", "SpmDriverHB"), ("orsirr.rua", "SpmDriverHB"),
                                         ("young4c.mtx", "SpmDriverMM"), ("mhd1280b.mtx", "SpmDriverMM"),
                                         ("lap", None)])
def test_nested_dissection_is_a_valid_fill_reducing_order(name, driver):
    if driver is None:
        spm = px.Spm.laplacian7(16, 16, 16)
    else:
        spm = px.Spm.read(getattr(px, driver), os.path.join(MAT, name))
    res = {}
    for ordering in (px.PastixOrderScotch, px.PastixOrderPersonal):
        iparm, dparm = px.initParam()
        iparm[px.IPARM_ORDERING] = ordering
        iparm[px.IPARM_VERBOSE] = px.PastixVerboseNot
        iparm[px.IPARM_FACTORIZATION] = px.PastixFactLU
        pd = px.init(iparm, dparm)
        px.task_analyze(pd, spm)
        perm, peri, rang, tree = px.order_export(pd)
        assert np.array_equal(np.sort(perm), np.arange(spm.n)) and np.array_equal(perm[peri], np.arange(spm.n))
        assert px.lib.pastixOrderCheck(px.lib.pastixOrderGet(pd)) == 0
        res[ordering] = dparm[px.DPARM_FACT_THFLOPS]
        px.finalize(pd)
    # less work than the natural order (PastixOrderPersonal without a
    # permutation) on the 3-D grid and the KKT matrix; no worse than 1.5x of
    # it on the small 2-D-like matrices, whose natural order is a good band
    if name in ("lap", "small.rsa"):
        assert res[px.PastixOrderScotch] < 0.5 * res[px.PastixOrderPersonal]
    else:
        assert res[px.PastixOrderScotch] < 1.5 * res[px.PastixOrderPersonal]


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("name,driver", [("young4c.mtx", "SpmDriverMM"), ("orsirr.rua", "SpmDriverHB"),
                                         ("mhd1280b.mtx", "SpmDriverMM")])
def test_symbolic_structure_holds_the_factor_for_any_ordering(name, driver, seed):
    """Block symbolic factorization on arbitrary permutations given without
    supernodes (fundamental supernodes, then fax): every nonzero of the dense
    Cholesky factor of the permuted pattern lies inside a blok.  Seed 0 is
    the natural order.  (A child cblk numbered just before its father used to
    lose its contribution.)"""
    import scipy.sparse as sp
    spm = px.Spm.read(getattr(px, driver), os.path.join(MAT, name))
    n = spm.n
    cp, rp, v = spm.arrays()
    A = sp.csc_matrix((np.ones(len(rp)), rp, cp), shape=(n, n))
    A = ((A + A.T) != 0).astype(float)
    perm = np.arange(n, dtype=np.int64) if seed == 0 else np.random.default_rng(seed).permutation(n).astype(np.int64)
    iparm, dparm = px.initParam()
    iparm[px.IPARM_ORDERING] = px.PastixOrderPersonal
    iparm[px.IPARM_VERBOSE] = px.PastixVerboseNot
    pd = px.init(iparm, dparm)
    o = px.pastix_order_t()
    px.lib.pastixOrderAlloc(ctypes.byref(o), n, 0)
    np.ctypeslib.as_array(o.permtab, shape=(n,))[:] = perm
    np.ctypeslib.as_array(o.peritab, shape=(n,))[perm] = np.arange(n)
    px._check("order", px.lib.pastix_subtask_order(pd, ctypes.byref(spm.s), ctypes.byref(o)))
    px.lib.pastixOrderExit(ctypes.byref(o))
    px.subtask_symbfact(pd)
    px.subtask_reordering(pd)
    final = px.order_export(pd)[0]
    cb, bk = px.symbol_export(pd)
    P = np.empty(n, dtype=np.int64)
    P[final] = np.arange(n)
    B = A[P][:, P].toarray() + np.eye(n) * 100
    L = np.abs(np.linalg.cholesky(B)) > 1e-12
    M = np.zeros((n, n), dtype=bool)
    for b in bk:
        M[b[0]:b[1] + 1, cb[b[2], 0]:cb[b[2], 1] + 1] = True
    assert not np.any(L & ~M)
    px.finalize(pd)
