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


def _header(name):
    return open(os.path.join(ROOT, "include", name)).read()


def test_python_constants_match_the_header():
    """Every MDSCTK_* value the binding mirrors equals the header's."""
    defs = {m.group(1): int(m.group(2))
            for m in re.finditer(r"#define MDSCTK_(\w+)\s+\(?(-?\d+)u?\)?", _header("mdsctk_knn.h"))}
    mirrored = {name: getattr(knn, name) for name in dir(knn)
                if re.fullmatch(r"(PREC|KNN|OPT|ERR|METRIC)_\w+", name)}
    assert mirrored, "no constants found in the binding"
    for name, value in mirrored.items():
        assert name in defs, name
        assert defs[name] == value, (name, defs[name], value)


def _struct_fields(text, name):
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    return [(t, f) for t, f in re.findall(r"(int32_t|int64_t|double)\s+(\w+);", body)]


@pytest.mark.parametrize("cname,pycls", [("mdsctk_knn_stats", "Stats"),
                                         ("mdsctk_knn_request_info", "RequestInfo"),
                                         ("mdsctk_knn_serve_stats", "ServeStats")])
def test_python_structs_match_the_header(cname, pycls):
    """The ctypes mirrors have the header's fields, types and order."""
    ctype = {"int32_t": ctypes.c_int32, "int64_t": ctypes.c_int64, "double": ctypes.c_double}
    want = [(f, ctype[t]) for t, f in _struct_fields(_header("mdsctk_knn.h"), cname)]
    got = list(getattr(knn, pycls)._fields_)
    assert want and got == want
