"""The C-ABI library loads and exports every entry point declared in include/smcf.h
(no device calls)."""
import ctypes
import os
import re

import pytest

import smc_inversion_fort_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "smcf.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b((?:smcf_\w+)|qsort_)\s*\(", txt)
    return sorted(set(names))


def test_header_matches_python_list():
    assert declared_symbols() == sorted(S.EXPORTS)


def test_library_exports_all_symbols():
    if not os.path.exists(S.LIB_PATH):
        pytest.skip("libsmcf.so not built")
    lib = ctypes.CDLL(S.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_qsort_matches_reference_golden(golden):
    if not os.path.exists(S.LIB_PATH):
        pytest.skip("libsmcf.so not built")
    import numpy as np
    g = golden("sort")
    i = 0
    while "c%d__arr" % i in g:
        idx = S.qsort(g["c%d__arr" % i], np.arange(1, len(g["c%d__arr" % i]) + 1))
        np.testing.assert_array_equal(idx, g["c%d__index" % i])
        i += 1
