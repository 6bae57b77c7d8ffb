"""CPU tests of the host side: the C-ABI library loads and exports every
declared entry point, argument checks that never touch the device, and the
Python API's geometry helpers against the oracle.  No GPU needed."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import rfp_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "curfp_amd", "libcurfp_amd.so")
HEADER = os.path.join(ROOT, "include", "curfp.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"curfpStatus_t\s+(curfp\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libcurfp_amd.so is not built (python curfp_amd/_build.py)")
    return ctypes.CDLL(LIB)


def test_header_declares_hot_path():
    syms = declared_symbols()
    for name in ["curfpCreate", "curfpDestroy", "curfpSetStream", "curfpGetStream",
                 "curfpSsfrk", "curfpSpftrf", "curfpSpftrs", "curfpSstrttf", "curfpSstfttr",
                 "curfpDsfrk", "curfpDpftrf", "curfpDpftrs", "curfpDstrttf", "curfpDstfttr",
                 "curfpSaddToDiagonal", "curfpDaddToDiagonal"]:
        assert name in syms


def test_library_exports_every_declared_symbol(lib):
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_null_handle_is_not_initialized(lib):
    """Reference src/curfp_internal.h:62-65: a null handle -> status 1."""
    one = ctypes.c_float(1.0)
    info = ctypes.c_int(0)
    assert lib.curfpSsfrk(None, 1, 1, 1, 4, 2, ctypes.byref(one), None, 2, ctypes.byref(one), None) == 1
    assert lib.curfpSpftrf(None, 1, 1, 4, None, ctypes.byref(info)) == 1
    assert lib.curfpSpftrs(None, 1, 1, 4, 1, None, None, 4) == 1
    assert lib.curfpSstrttf(None, 1, 1, 4, None, 4, None) == 1
    assert lib.curfpSetStream(None, None) == 1
    assert lib.curfpCreate(None) == 3


def test_fake_handle_argument_checks(lib):
    """Argument checks run before any device work (reference
    src/curfp_ssfrk.cpp:149-150, src/curfp_spftrf.cpp:191-192,
    src/curfp_spftrs.cpp:226, src/curfp_strttf.cu:430).  A dummy handle
    object is enough because these paths return before using it."""
    dummy = ctypes.create_string_buffer(64)
    h = ctypes.cast(dummy, ctypes.c_void_p)
    one = ctypes.c_float(1.0)
    info = ctypes.c_int(0)
    assert lib.curfpSsfrk(h, 1, 1, 1, -1, 2, ctypes.byref(one), None, 2, ctypes.byref(one), None) == 3
    assert lib.curfpSsfrk(h, 1, 1, 1, 4, -1, ctypes.byref(one), None, 2, ctypes.byref(one), None) == 3
    assert lib.curfpSsfrk(h, 1, 1, 1, 4, 2, ctypes.byref(one), None, 0, ctypes.byref(one), None) == 3
    assert lib.curfpSsfrk(h, 1, 1, 1, 4, 2, None, None, 2, ctypes.byref(one), None) == 3
    # n == 0 and the alpha=0 / beta=1 no-op are successes that touch nothing
    assert lib.curfpSsfrk(h, 1, 1, 1, 0, 2, ctypes.byref(one), None, 2, ctypes.byref(one), None) == 0
    zero = ctypes.c_float(0.0)
    assert lib.curfpSsfrk(h, 1, 1, 1, 4, 2, ctypes.byref(zero), None, 2, ctypes.byref(one), None) == 0
    assert lib.curfpSpftrf(h, 1, 1, -1, None, ctypes.byref(info)) == 3
    assert lib.curfpSpftrf(h, 1, 1, 4, None, None) == 3
    info.value = 7
    assert lib.curfpSpftrf(h, 1, 1, 0, None, ctypes.byref(info)) == 0 and info.value == 0
    assert lib.curfpSpftrs(h, 1, 1, -1, 1, None, None, 4) == 3
    assert lib.curfpSpftrs(h, 1, 1, 4, -1, None, None, 4) == 3
    assert lib.curfpSpftrs(h, 1, 1, 4, 1, None, None, 0) == 3
    assert lib.curfpSpftrs(h, 1, 1, 4, 0, None, None, 4) == 0
    assert lib.curfpSstrttf(h, 1, 1, 4, None, 3, None) == 3
    assert lib.curfpSstfttr(h, 1, 1, -1, None, None, 4) == 3
    assert lib.curfpSsfrk(h, 2, 1, 1, 4, 2, ctypes.byref(one), None, 2, ctypes.byref(one), None) == 3


def test_python_geometry_matches_oracle():
    import torch

    import curfp_amd as curfp

    for n in [1, 2, 3, 6, 7, 64, 65, 100, 101]:
        for t in "NT":
            for u in "LU":
                n1, n2, ld, v11, v21, v22 = curfp.rfp_blocks(n, t, u)
                assert (n1, n2, ld, v11, v21, v22) == O.views(n, t, u)
                d = curfp.rfp_diag_indices(n, t, u, device="cpu").numpy()
                assert np.array_equal(d, O.diag_indices(n, t, u))
    assert curfp._n_from_rfp(65536 * 65537 // 2) == 65536
    with pytest.raises(ValueError):
        curfp._n_from_rfp(11)
    with pytest.raises(ValueError):
        curfp._op("X")
    with pytest.raises(ValueError):
        curfp._fill("Q")
    with pytest.raises(ValueError):
        curfp.ssfrk(torch.zeros(4, 2), torch.zeros(10))  # not a device tensor


# the public functions of the reference's Python package (curfp/__init__.py)
REFERENCE_API = [
    "set_stream", "rfp_diag_indices", "add_to_diagonal", "ssfrk_raw", "spftrf_raw", "spftrs_raw", "spftri_raw",
    "ssfrk", "spftrf", "spftrs", "ssfmv_raw", "ssfmv", "spftri", "strttf_raw", "stfttr_raw", "strttf", "stfttr",
    "slansf_raw", "slansf", "spfcon_raw", "spfcon", "spftrf_rcond", "ssfr", "ssfr_raw", "ssfr2", "ssfr2_raw",
    "ssfr2k", "ssfr2k_raw", "ssfmm", "ssfmm_raw", "dstrttf", "dstfttr", "dstrttf_raw", "dstfttr_raw", "dsfrk",
    "dsfrk_raw", "dpftrf", "dpftrf_raw", "dpftrs", "dpftrs_raw", "dpftri", "dpftri_raw", "dsfmv", "dsfmv_raw",
    "dlansf", "dlansf_raw", "dpfcon", "dpfcon_raw", "dpftrf_rcond", "dsfr", "dsfr_raw", "dsfr2", "dsfr2_raw",
    "dsfr2k", "dsfr2k_raw", "dsfmm", "dsfmm_raw",
]


def test_reference_api_is_complete():
    import curfp_amd as curfp

    missing = [f for f in REFERENCE_API if not callable(getattr(curfp, f, None))]
    assert not missing, missing


REF_EXT = "/root/reference/curfp/_ext.cpp"


def _reference_bindings():
    """(functions {name: [keyword names]}, module constants, Handle methods)
    as the reference module declares them (curfp/_ext.cpp:672-878)."""
    text = open(REF_EXT).read()
    text = text[text.index("PYBIND11_MODULE"):]
    funcs = {}
    for m in re.finditer(r'm\.def\("(\w+)",\s*&\w+\s*,(.*?)\);', text, re.S):
        funcs[m.group(1)] = re.findall(r'py::arg\("(\w+)"\)', m.group(2))
    attrs = re.findall(r'm\.attr\("(\w+)"\)', text)
    methods = dict((m.group(1), re.findall(r'py::arg\("(\w+)"\)', m.group(2) or ""))
                   for m in re.finditer(r'\.def\("(\w+)",\s*&Handle::\w+(.*?)(?=\n\s*\.def|\);)', text, re.S))
    return funcs, attrs, methods


def _keywords(fn):
    sig = fn.__doc__.splitlines()[0]
    return re.findall(r"(\w+):", sig[sig.index("(") + 1:sig.rindex(")")])


@pytest.mark.skipif(not os.path.exists(REF_EXT), reason="the reference sources are not present")
def test_extension_matches_reference_bindings():
    """Every function of the reference's _curfp_C exists here under the same
    name with the same keyword names in the same order, and every module
    constant has the same value (curfp/_ext.cpp:684-878)."""
    from curfp_amd import _curfp_C as ext

    funcs, attrs, methods = _reference_bindings()
    assert len(funcs) >= 26 and len(attrs) == 9
    missing = [f for f in funcs if not hasattr(ext, f)]
    assert not missing, missing
    mismatched = {f: (_keywords(getattr(ext, f)), kw) for f, kw in funcs.items() if _keywords(getattr(ext, f)) != kw}
    assert not mismatched, mismatched
    values = {"OP_N": 0, "OP_T": 1, "FILL_LOWER": 0, "FILL_UPPER": 1, "NORM_MAX": 0, "NORM_ONE": 1, "NORM_FRO": 2,
              "SIDE_LEFT": 0, "SIDE_RIGHT": 1}
    assert sorted(attrs) == sorted(values)
    assert {a: getattr(ext, a) for a in attrs} == values
    for name, kw in methods.items():
        assert hasattr(ext.Handle, name), name
        if kw:
            assert _keywords(getattr(ext.Handle, name))[1:] == kw, name


def test_import_as_reference_name():
    """``import curfp`` gives this package, with the extension at
    curfp._curfp_C (the names the reference's code imports)."""
    import curfp
    import curfp_amd

    assert curfp is curfp_amd
    from curfp import _curfp_C, NORM_FRO, SIDE_RIGHT  # noqa: F401
    assert _curfp_C is curfp_amd._curfp_C
