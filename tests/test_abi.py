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


def test_ctypes_mirrors_match_the_c_layout(tmp_path):
    """every field offset and the size of the Python mirrors of the C-ABI structs equal the C compiler's (a small C
    program compiled here prints them)"""
    import subprocess
    import g3d_amd
    mirrors = {"g3d_axis": g3d_amd.Axis, "g3d_plane_wave": g3d_amd.PlaneWave, "g3d_push_ctx": g3d_amd.PushCtx,
               "g3d_parts": g3d_amd.Parts, "g3d_stream_ctx": g3d_amd.StreamCtx, "g3d_spec_ctx": g3d_amd.SpecCtx,
               "g3d_config": g3d_amd.Config, "g3d_sim_info": g3d_amd.SimInfo}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % HEADER, "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        a, b, v = line.split()
        got[(a, b)] = int(v)
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
