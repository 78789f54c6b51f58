"""The C-ABI library loads and exports every function include/mlhip.h
declares (no GPU calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "mlhip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(ml_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(ml):
    lib = ml.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_header(ml):
    assert sorted(ml.SIGNATURES) == declared_functions()


def test_layout_helpers_without_gpu(ml):
    lib = ml.lib()
    n = (ctypes.c_int * 3)(64, 32, 1)
    assert lib.ml_cell_count(2, n) == 66 * 34
    assert lib.ml_face_count(2, n, 0) == 65 * 32
    assert lib.ml_face_count(2, n, 1) == 64 * 33
    assert lib.ml_mg_depth(2, n) == 5          # 64x32 -> ... -> 4x2
    n3 = (ctypes.c_int * 3)(16, 16, 8)
    assert lib.ml_cell_count(3, n3) == 18 * 18 * 10
    assert lib.ml_mg_depth(3, n3) == 3         # 16x16x8 -> 8x8x4 -> 4x4x2
    o = ml.default_opts()
    assert (o.nu1, o.nu2, o.bottom, o.max_iter) == (1, 4, 40, 10)
    assert o.eps == o.hang == o.norm_thresh == 1e-10


def test_io_library_exports_every_declared_symbol():
    import importlib
    b = importlib.import_module("mushy-layer_amd.build")
    cio = importlib.import_module("mushy-layer_amd.chombo_io")
    if not os.path.exists(cio.IO_LIB) and b.build_io() is None:
        import pytest
        pytest.skip("no HDF5 library to build libmlio.so")
    src = open(os.path.join(ROOT, "include", "mlio.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(ml_h5_[a-z0-9_]+)\s*\(", src)))
    lib = cio.lib()
    assert [n for n in names if not hasattr(lib, n)] == []
    assert sorted(cio._SIG) == names
