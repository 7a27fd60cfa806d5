"""CPU-side checks of the C ABI: the library loads without a GPU, exports
every function include/filtergs_amd.h declares, and its host-only queries
behave."""
import os
import re

from filtergs_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "filtergs_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return set(re.findall(r"\b(fgs_[a-z0-9_]+)\s*\(", src))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _declared()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(_lib.EXPORTED)


def test_host_queries():
    lib = _lib.load()
    assert lib.fgs_abi_version() == 1
    assert lib.fgs_geom_bytes(0) > 0
    assert lib.fgs_geom_bytes(2000) > lib.fgs_geom_bytes(1000)
    assert lib.fgs_image_bytes(1920, 1080) >= 1920 * 1080 * 8 + 120 * 68 * 8
    assert lib.fgs_binning_bytes(10 ** 6, 1920, 1080) >= 16 * 10 ** 6
    assert lib.fgs_backward_ws_bytes(1000) >= 12000


def test_ogfs_constant_matches_oracle(orc):
    lib = _lib.load()
    for es, em in ((0.2, 0.2), (1.17, 0.2), (0.0, 0.5), (3.0, 0.9)):
        assert lib.fgs_ogfs_log_lambda(es, em) == orc.ogfs_log_lambda(es, em)


def test_null_arguments_fail_cleanly():
    import ctypes
    lib = _lib.load()
    rc = lib.fgs_forward_geometry(None, None, None, None, None, None)
    assert rc == 1
    assert b"null" in lib.fgs_last_error()
    rc = lib.fgs_mark_visible(-1, None, None, None, None, None)
    assert rc == 1
    del ctypes


def test_traverse_workspace_query():
    lib = _lib.load()
    N, C, R = 100000, 4, 300
    assert lib.fgs_tree_traverse_ws_bytes(N, C, R) >= 3 * 4 * (N + R) + C * (N + R)
    assert lib.fgs_tree_traverse_ws_bytes(N, 8, R) > lib.fgs_tree_traverse_ws_bytes(N, 4, R)
    assert lib.fgs_tree_traverse_ws_bytes(-1, C, R) == 0
    assert lib.fgs_tree_traverse_ws_bytes(N, 0, R) == 0
