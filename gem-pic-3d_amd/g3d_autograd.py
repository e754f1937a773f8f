"""Gradients through the march: a torch.autograd.Function over the C ABI of libg3d.so.

forward runs g3d_march (the per-mode march of every slice, front to rear, Fourier y) on copies of its inputs;
backward runs g3d_march_adjoint, the hand-written transpose of the same update, rear to front.  Tensors are
complex128 on the GPU, one row per slice (ns rows for the fields, ns + 1 for the currents, nk = nx ny modes in the
reference's layout j nx + ix).  The geometry (k vectors, half-cell lengths, current weights) is taken from an
engine built on the same deck.  No counterpart in the reference.

torch runs its kernels on its current stream and the library on its own, so each call synchronises the device
before and after the library call.
"""
import ctypes as C

import torch

import g3d_amd

_FIELDS = ("Rx", "Ry", "Ez", "Bz", "Lx", "Ly")
_CURRENTS = ("j0x", "j0y", "jkx", "jky", "jkz")


class MarchCtx(C.Structure):
    """g3d_march_ctx (include/g3d.h)"""
    _fields_ = [(n, C.c_void_p) for n in _FIELDS + ("bnd",) + _CURRENTS + ("bkRx", "bkRy", "bkEz", "bkBz", "kx", "ky",
                                                                                "iD", "iA", "hcg", "wlo", "whi")] + \
               [("D", C.c_double), ("a", C.c_int), ("b", C.c_int), ("N", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
                ("nk", C.c_long), ("scp", C.c_void_p)]


class MarchAdjCtx(C.Structure):
    """g3d_march_adj_ctx (include/g3d.h)"""
    _fields_ = [(n, C.c_void_p) for n in _FIELDS + ("up",) + _CURRENTS + ("kx", "ky", "hcg", "wlo", "whi")] + \
               [("D", C.c_double), ("a", C.c_int), ("N", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nk", C.c_long)]


class MarchGeometry:
    """The march constants of a deck (Fourier y), on the GPU as float64 tensors."""

    def __init__(self, sim, D):
        info = sim.info()
        if info["N"] != info["ns"]:
            raise ValueError("the march geometry needs the whole window on this GPU (a = 0)")
        dev = torch.device("cuda")
        self.N, self.ns, self.nk = int(info["N"]), int(info["ns"]), int(info["nk"])
        self.D = float(D)
        self.tables = {n: torch.from_numpy(sim.array(n)).to(dev) for n in ("kx", "ky", "hcg", "wlo", "whi")}
        self.nx = int(self.tables["kx"].numel())
        self.ny = int(self.tables["ky"].numel())
        if self.nx * self.ny != self.nk:
            raise ValueError("free-hand y has no Fourier march geometry")

    def _common(self, c):
        for n, t in self.tables.items():
            setattr(c, n, t.data_ptr())
        c.D, c.a, c.N, c.nx, c.ny, c.nk = self.D, 0, self.N, self.nx, self.ny, self.nk


def _check_planes(t, rows, nk, what):
    if t.dtype != torch.complex128 or not t.is_cuda or tuple(t.shape) != (rows, nk):
        raise ValueError("%s: expected a complex128 GPU tensor of shape (%d, %d)" % (what, rows, nk))


def march(geom, old, currents):
    """old: (R_x, R_y, E_z, B_z) at level s0, each (ns, nk); currents: (j0x, j0y, jkx, jky, jkz), each (ns + 1, nk).
    Returns the six fields (R_x, R_y, E_z, B_z, L_x, L_y) of every slice at level s1, differentiable."""
    return MarchFunction.apply(geom, *old, *currents)


class MarchFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, geom, Rx, Ry, Ez, Bz, j0x, j0y, jkx, jky, jkz):
        ns, nk = geom.ns, geom.nk
        for t, n in zip((Rx, Ry, Ez, Bz), _FIELDS):
            _check_planes(t, ns, nk, n)
        cur = [t.contiguous() for t in (j0x, j0y, jkx, jky, jkz)]
        for t, n in zip(cur, _CURRENTS):
            _check_planes(t, ns + 1, nk, n)
        out = [t.contiguous().clone() for t in (Rx, Ry, Ez, Bz)]
        out += [torch.zeros_like(out[0]), torch.zeros_like(out[0])]
        c = MarchCtx()
        for n, t in zip(_FIELDS, out):
            setattr(c, n, t.data_ptr())
        for n, t in zip(_CURRENTS, cur):
            setattr(c, n, t.data_ptr())
        geom._common(c)
        c.b = geom.N
        lib = g3d_amd.lib()
        torch.cuda.synchronize()
        g3d_amd.check(lib.g3d_march(C.byref(c), 0, geom.N))
        g3d_amd.check(lib.g3d_sync())
        ctx.geom = geom
        return tuple(out)

    @staticmethod
    def backward(ctx, *grads):
        geom = ctx.geom
        ns, nk, dev = geom.ns, geom.nk, torch.device("cuda")
        bars = [(g.contiguous().clone() if g is not None else torch.zeros(ns, nk, dtype=torch.complex128, device=dev))
                for g in grads]
        cur = [torch.zeros(ns + 1, nk, dtype=torch.complex128, device=dev) for _ in _CURRENTS]
        c = MarchAdjCtx()
        for n, t in zip(_FIELDS, bars):
            setattr(c, n, t.data_ptr())
        for n, t in zip(_CURRENTS, cur):
            setattr(c, n, t.data_ptr())
        c.up = None
        geom._common(c)
        lib = g3d_amd.lib()
        torch.cuda.synchronize()
        g3d_amd.check(lib.g3d_march_adjoint(C.byref(c), 0, geom.N))
        g3d_amd.check(lib.g3d_sync())
        return (None, bars[0], bars[1], bars[2], bars[3], *cur)
