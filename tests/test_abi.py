"""The C ABI library loads and exports every symbol the public headers declare."""
import ctypes
import glob
import os
import re

import pytest

from mdsctk_amd import knn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        names.update(re.findall(r"^\s*(?:const\s+char\s*\*|int|void|double)\s+(mdsctk_\w+)\s*\(", text, re.M))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_functions()
    for must in ["mdsctk_knn_create", "mdsctk_knn_set_reference", "mdsctk_knn_query",
                 "mdsctk_knn_destroy", "mdsctk_xtc_read", "mdsctk_topology_read"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    L = knn.load()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_no_device_errors_are_codes_not_crashes():
    L = knn.load()
    h = ctypes.c_void_p()
    assert L.mdsctk_knn_create(ctypes.byref(h), 0, 99) == knn.ERR_ARG
    assert L.mdsctk_knn_create(None, 0, 1) == knn.ERR_ARG
    assert L.mdsctk_knn_destroy(None) == knn.ERR_ARG
    assert L.mdsctk_knn_last_error(None) == b"null context"
    assert L.mdsctk_knn_query(None, None, 0, 1, 0, None, None) == knn.ERR_ARG
