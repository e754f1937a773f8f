"""Test-side wrapper of oracle/_ref/libg3dref.so: the reference's own Simulation (its CPU build, compiled from
/root/reference by oracle/Makefile) behind the same Python front as the engine (g3d_amd.Sim).

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline use it; it is the checker, never the product.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "gem-pic-3d_amd"))
import g3d_amd  # noqa: E402

REF_PATH = os.path.join(HERE, "_ref", "libg3dref.so")
_lib = None


def available():
    return os.path.exists(REF_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise ImportError("oracle/_ref/libg3dref.so is not built (make -C oracle ref, needs the reference tree)")
        L = C.CDLL(REF_PATH)
        P = C.c_void_p
        sig = {
            "ref_last_error": ([], C.c_char_p), "ref_set_threads": ([C.c_int], None),
            "ref_sim_create": ([P], P), "ref_sim_destroy": ([P], None), "ref_sim_info": ([P, P], C.c_int),
            "ref_sim_call": ([P, C.c_char_p, C.POINTER(C.c_double), C.c_int], C.c_int),
            "ref_sim_output": ([P, C.c_char_p, C.c_char_p, C.c_long, C.c_double, C.POINTER(C.c_double)], C.c_int),
            "ref_sim_array_size": ([P, C.c_char_p, C.POINTER(C.c_size_t)], C.c_int),
            "ref_sim_get": ([P, C.c_char_p, P, C.c_size_t], C.c_int),
            "ref_sim_set": ([P, C.c_char_p, P, C.c_size_t], C.c_int),
            "ref_sim_set_particles": ([P, P, C.c_long], C.c_int), "ref_sim_set_joining": ([P, P, C.c_long], C.c_int),
            "ref_sim_get_joining": ([P, P, C.c_long, C.POINTER(C.c_long)], C.c_int),
            "ref_fft_many": ([C.c_int, C.c_int, C.c_int, P, C.c_int, C.c_int], C.c_int),
            "ref_adk_levels": ([P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, P], C.c_int),
            "ref_adk_rate_SI": ([C.c_double, C.c_int, C.c_double], C.c_double),
            "ref_mix64": ([C.c_ulonglong], C.c_ulonglong), "ref_uniform01": ([C.c_ulonglong], C.c_double),
            "ref_key_step": ([C.c_ulonglong, C.c_long], C.c_ulonglong),
            "ref_key_slice": ([C.c_ulonglong, C.c_int, C.c_int], C.c_ulonglong),
            "ref_key_particle": ([C.c_ulonglong, C.c_ulonglong], C.c_ulonglong),
            "ref_deck_config": ([C.c_char_p, C.POINTER(C.c_char_p), C.c_int, P, P, C.c_int, P, C.c_int, P, C.c_char_p,
                                 C.c_int], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise RuntimeError("reference: " + lib().ref_last_error().decode())


def set_threads(n):
    lib().ref_set_threads(int(n))


def deck_config(path, overrides=(), zcap=1 << 16, ycap=1 << 14):
    """the reference's own reading of a deck file (Deck::read, overrides, Config::from_deck): (g3d_config, zeta,
    ynode, output settings as a dict)"""
    cfg = g3d_amd.Config()
    zeta, ynode, out = np.zeros(zcap), np.zeros(ycap), np.zeros(10)
    d = C.create_string_buffer(4096)
    ov = (C.c_char_p * max(1, len(overrides)))(*[o.encode() for o in overrides])
    check(lib().ref_deck_config(str(path).encode(), ov, len(overrides), C.byref(cfg), zeta.ctypes.data, zcap,
                                ynode.ctypes.data, ycap, out.ctypes.data, d, len(d)))
    keys = ("every", "beam_every", "fields", "beams", "axis", "particles", "frames", "slice_length",
            "checkpoint_every", "walltime_hours")
    o = {k: float(v) for k, v in zip(keys, out)}
    o["dir"] = d.value.decode()
    return cfg, zeta[:cfg.nzeta].copy(), ynode[:cfg.ny + 1].copy(), o


class RefSim(g3d_amd.SimBase):
    def __init__(self, deck):
        self._deck = deck
        h = lib().ref_sim_create(deck.ptr())
        if not h:
            raise RuntimeError("reference: " + lib().ref_last_error().decode())
        self.h = h

    def close(self):
        if self.h:
            lib().ref_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self):
        self.call("init")

    def step(self):
        self.call("step")

    def call(self, op, *args):
        a = (C.c_double * max(1, len(args)))(*args)
        check(lib().ref_sim_call(self.h, op.encode(), a, len(args)))

    def _size(self, name):
        b = C.c_size_t()
        check(lib().ref_sim_array_size(self.h, name.encode(), C.byref(b)))
        return b.value

    def _get(self, name, out):
        check(lib().ref_sim_get(self.h, name.encode(), out.ctypes.data_as(C.c_void_p), out.nbytes))

    def set(self, name, arr):
        arr = np.ascontiguousarray(arr)
        check(lib().ref_sim_set(self.h, name.encode(), arr.ctypes.data_as(C.c_void_p), arr.nbytes))

    def _output(self, what, path="", arg=0, darg=0.0):
        out = C.c_double()
        check(lib().ref_sim_output(self.h, what.encode(), str(path).encode(), int(arg), float(darg), C.byref(out)))
        return out.value

    def write_bunches(self, path, header=False):
        self._output("bunches", path, int(header))

    def write_slices(self, path, slice_length):
        self._output("slices", path, 0, slice_length)

    def diagnostics(self, path, header=False):
        self._output("diag", path, int(header))

    def field_energy(self):
        return self._output("energy")

    def dump_particles(self, path):
        self._output("particles", path)

    def write_frame(self, path, comps=255):
        self._output("frame", path, comps)

    def write_axis(self, path):
        self._output("axis", path)

    def write_checkpoint(self, directory):
        self._output("write_ckpt", directory)

    def read_checkpoint(self, directory, step):
        self._output("read_ckpt", directory, step)

    def set_particles(self, p):
        p = np.ascontiguousarray(p, dtype=g3d_amd.PARTICLE)
        check(lib().ref_sim_set_particles(self.h, p.ctypes.data_as(C.c_void_p), len(p)))

    def set_joining(self, p):
        p = np.ascontiguousarray(p, dtype=g3d_amd.PARTICLE)
        check(lib().ref_sim_set_joining(self.h, p.ctypes.data_as(C.c_void_p), len(p)))

    def joining(self, cap=1 << 22):
        out = np.empty(cap, dtype=g3d_amd.PARTICLE)
        n = C.c_long()
        check(lib().ref_sim_get_joining(self.h, out.ctypes.data_as(C.c_void_p), cap, C.byref(n)))
        return out[:n.value].copy()

    def _info(self, i):
        check(lib().ref_sim_info(self.h, C.byref(i)))


def fft_many(nx, ny, two_d, data, inverse):
    """The reference's FftMany::run on a host array of shape (count, ny, nx), in place."""
    a = np.array(data, dtype=np.complex128, copy=True, order="C")
    check(lib().ref_fft_many(nx, ny, int(two_d), a.ctypes.data_as(C.c_void_p), a.shape[0], int(inverse)))
    return a
