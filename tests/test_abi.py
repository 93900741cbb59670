"""The C ABI library loads and exports every function include/giem2g_amd.h
declares (no compute calls: this runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "giem2g_amd.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(giem2g_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_header_declares_reference_entry_points():
    names = declared_functions()
    for n in ["giem2g_calc_data_sizes", "giem2g_prepare_ie_kernel", "giem2g_calc_ie_kernel",
              "giem2g_calc_fft_of_ie_kernel", "giem2g_set_anomaly_conductivity",
              "giem2g_apply_ie_operator", "giem2g_set_logger", "giem2g_print_stats",
              "giem2g_delete_ie_operator"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    from giem2g_amd import lib as L
    if not L.available():
        pytest.skip("libgiem2g_amd.so not built")
    so = C.CDLL(L.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(so, n)]
    assert not missing, missing
    assert set(L.REFERENCE_SYMBOLS + L.EXTENDED_SYMBOLS) == set(declared_functions())


def test_product_package_does_not_import_oracle():
    pkg = os.path.join(ROOT, "giem2g_amd")
    for dp, _, fs in os.walk(pkg):
        for f in fs:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                txt = open(os.path.join(dp, f)).read()
                assert "oracle" not in txt.lower() or f == "__init__.py" and "oracle" not in txt, f
