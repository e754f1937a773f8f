"""Build recipe of the gfx950 library libg3d.so (in-tree, next to the sources).

hipcc compiles every source for gfx950 with:
  -ffp-contract=off                          no fused multiply-add, so that the FP64 kernels round like the
                                             reference's g++ -O3 x86-64 build (which has no FMA)
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
Objects go to _build/; the library to gem-pic-3d_amd/lib/libg3d.so.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libg3d.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["g3d_fft.hip", "g3d_march.hip", "g3d_particles.hip", "g3d_fields.hip", "g3d_plasma.hip",
           "g3d_freehand.hip", "g3d_deposit.hip", "g3d_mixed.hip", "g3d_adjoint.hip", "g3d_spush.hip", "g3d_bpush.hip", "g3d_kabi.hip", "g3d_static.hip", "g3d_grad.hip",
           "g3d_engine.cpp", "g3d_output.cpp", "g3d_grad.cpp", "g3d_abi.cpp"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-result"]


def _compile(src):
    obj = os.path.join(OBJ, src + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
