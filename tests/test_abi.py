"""The C-ABI library loads and exports every function its headers declare."""
import ctypes
import os
import re

import pytest

import pastix_amd as px

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = ["include/pastix.h", "include/pastix/order.h", "include/spm.h",
           "include/pastix_amd_device.h", "include/pastix_amd_engine.h", "include/pastix_amd_models.h"]


def declared_functions(path):
    text = open(os.path.join(ROOT, path)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"#.*", "", text)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|void|double|size_t|char\s*\*|void\s*\*|pastix_model_t)\s*\**\s*"
                       r"([A-Za-z_]\w*)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize("header", HEADERS)
def test_every_declared_function_is_exported(header):
    names = declared_functions(header)
    assert names, header
    missing = [n for n in names if not hasattr(px.lib, n)]
    assert not missing, "%s: not exported: %s" % (header, missing)


def test_no_torch_or_hip_types_in_the_c_abi():
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, h)).read(), flags=re.S)
        for bad in ("hipStream_t", "torch", "at::", "std::"):
            assert bad not in text, (h, bad)


def test_enum_values_match_the_reference_abi():
    # values fixed by the reference interface (include/pastix/api.h)
    assert px.IPARM_SIZE == 80 and px.DPARM_SIZE == 26
    assert px.IPARM_FACTORIZATION == 43 and px.IPARM_MIXED == 71
    assert px.DPARM_FACT_THFLOPS == 12 and px.DPARM_A_NORM == 23
    assert px.PastixFactPOTRF == 0 and px.PastixFactGETRF == 2 and px.PastixFactSYTRF == 1
    assert px.PASTIX_ERR_BADPARAMETER == 7
    assert px.PastixDouble == 3 and px.PastixFloat == 2 and px.PastixComplex64 == 5
    assert px.PastixNoTrans == 111 and px.PastixLower == 122 and px.PastixDirForward == 391


def test_step_order_errors_without_gpu():
    iparm, dparm = px.initParam()
    pd = px.init(iparm, dparm)
    spm = px.Spm.laplacian7(3, 3, 3)
    assert px.lib.pastix_task_numfact(pd, ctypes.byref(spm.s)) == px.PASTIX_ERR_BADPARAMETER
    x = spm.gen_rhs(1)[1]
    assert px.lib.pastix_task_solve(pd, spm.n, 1, x.ctypes.data_as(ctypes.c_void_p), spm.n) == \
        px.PASTIX_ERR_BADPARAMETER
    assert px.lib.pastix_subtask_symbfact(pd) == px.PASTIX_ERR_BADPARAMETER
    assert px.lib.pastix_subtask_blend(pd) == px.PASTIX_ERR_BADPARAMETER
    px.finalize(pd)


def test_scotch_and_metis_orderings_run_the_built_in_nested_dissection():
    for ordering in (px.PastixOrderScotch, px.PastixOrderMetis):
        iparm, dparm = px.initParam()
        iparm[px.IPARM_ORDERING] = ordering
        pd = px.init(iparm, dparm)
        spm = px.Spm.laplacian7(3, 3, 3)
        assert px.lib.pastix_task_analyze(pd, ctypes.byref(spm.s)) == px.PASTIX_SUCCESS
        px.finalize(pd)


def _reference_prototypes():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    try:
        import make_reference_api
    finally:
        sys.path.pop(0)
    ours = {}
    for h in ("include/pastix.h", "include/pastix/order.h"):
        ours.update(make_reference_api.prototypes(open(os.path.join(ROOT, h)).read()))
    ref = {}
    with open(os.path.join(ROOT, "tests", "golden", "reference_api.txt")) as f:
        for line in f:
            line = line.strip()
            if line:
                ref[line.split("|")[0]] = line
    return ref, ours


def test_public_prototypes_match_the_reference_headers():
    # every prototype of the reference's include/pastix.h and
    # include/pastix/order.h, with the same return and parameter types
    ref, ours = _reference_prototypes()
    assert len(ref) >= 45
    diff = ["%s != %s" % (ours.get(n), v) for n, v in sorted(ref.items()) if ours.get(n) != v]
    assert not diff, diff


def test_reference_functions_are_exported():
    ref, _ = _reference_prototypes()
    missing = [n for n in ref if not hasattr(px.lib, n)]
    assert not missing, missing
