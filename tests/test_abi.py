"""The C-ABI library (gem-pic-3d_amd/lib/libg3d.so) loads without a GPU and exports every function include/g3d.h
declares.  No compute call is made here."""
import ctypes
import os
import re
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "g3d.h")
sys.path.insert(0, os.path.join(ROOT, "gem-pic-3d_amd"))


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"//[^\n]*", "", src)
    src = re.sub(r"^\s*#[^\n]*", "", src, flags=re.M)
    return sorted(set(re.findall(r"\b(g3d_\w+)\s*\(", src)))


def library():
    import g3d_amd
    if not os.path.exists(g3d_amd.LIB_PATH):
        import build
        build.build(verbose=False)
    return ctypes.CDLL(g3d_amd.LIB_PATH)


def test_header_declares_the_march_interface():
    names = declared_functions()
    for n in ("g3d_march", "g3d_march_slice", "g3d_fft_plan", "g3d_fft_exec", "g3d_sim_create", "g3d_sim_step"):
        assert n in names, n


def test_library_exports_every_declared_function():
    lib = library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings():
    lib = library()
    lib.g3d_status_string.restype = ctypes.c_char_p
    lib.g3d_status_string.argtypes = [ctypes.c_int]
    assert lib.g3d_status_string(0)
