"""Gradients through the march: torch.autograd.Functions over the C ABI of libg3d.so.  No counterpart in the
reference.

MarchFunction: forward runs g3d_march (the per-mode march of every slice, front to rear, Fourier y) on copies of
its inputs; backward runs g3d_march_adjoint, the hand-written transpose of the same update, rear to front.  Tensors
are complex128 on the GPU, one row per slice (ns rows for the fields, ns + 1 for the currents, nk = nx ny modes in
the reference's layout j nx + ix).  The geometry (k vectors, half-cell lengths, current weights) is taken from an
engine built on the same deck.

TscDeposit / TscGather: the TSC deposit of particle values onto a periodic plane and the gather from it, each the
other's adjoint, with the derivatives in the particle positions.  BorisFunction: the Boris rotation of the
streaming push with its hand-written vector-Jacobian product.  All float64 on the GPU.

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


# ---------------------------------------------------------------------------------------------- particle pieces
class TscCtx(C.Structure):
    """g3d_tsc_ctx (include/g3d.h)"""
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("np", C.c_long), ("nx", C.c_int), ("ny", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double)]


class BorisArgs(C.Structure):
    """g3d_boris_args (include/g3d.h)"""
    _fields_ = [(n, C.c_void_p) for n in ("ux", "uy", "uz", "Ex", "Ey", "Ez", "Bx", "By", "Bz", "qm", "dt")] + \
               [("np", C.c_long)]


class BorisAdj(C.Structure):
    """g3d_boris_adj (include/g3d.h)"""
    _fields_ = [(n, C.c_void_p) for n in ("ubx", "uby", "ubz", "gux", "guy", "guz", "gEx", "gEy", "gEz", "gBx",
                                          "gBy", "gBz", "gqm", "gdt")]


def _f64(t, what):
    if t.dtype != torch.float64 or not t.is_cuda or t.dim() != 1:
        raise ValueError("%s: expected a 1-D float64 GPU tensor" % what)
    return t.contiguous()


def _tsc_ctx(x, y, nx, ny, dx, dy):
    c = TscCtx()
    c.x, c.y, c.np, c.nx, c.ny, c.dx, c.dy = x.data_ptr(), y.data_ptr(), x.numel(), nx, ny, dx, dy
    return c


def _call(f, *args):
    torch.cuda.synchronize()
    g3d_amd.check(f(*args))
    g3d_amd.check(g3d_amd.lib().g3d_sync())


def tsc_deposit_raw(x, y, q, nx, ny, dx, dy):
    plane = torch.zeros(ny * nx, dtype=torch.float64, device=x.device)
    c = _tsc_ctx(x, y, nx, ny, dx, dy)
    _call(g3d_amd.lib().g3d_tsc_deposit, C.byref(c), C.c_void_p(q.data_ptr()), C.c_void_p(plane.data_ptr()))
    return plane


def tsc_gather_raw(x, y, plane, nx, ny, dx, dy, grad=False):
    out = torch.empty_like(x)
    gx = torch.empty_like(x) if grad else None
    gy = torch.empty_like(x) if grad else None
    c = _tsc_ctx(x, y, nx, ny, dx, dy)
    _call(g3d_amd.lib().g3d_tsc_gather, C.byref(c), C.c_void_p(plane.data_ptr()), C.c_void_p(out.data_ptr()),
          C.c_void_p(gx.data_ptr() if grad else None), C.c_void_p(gy.data_ptr() if grad else None))
    return (out, gx, gy) if grad else out


class TscDeposit(torch.autograd.Function):
    """plane[j nx + i] = sum_p W(x_p, y_p; i, j) q_p (periodic TSC)"""
    @staticmethod
    def forward(ctx, x, y, q, nx, ny, dx, dy):
        x, y, q = _f64(x, "x"), _f64(y, "y"), _f64(q, "q")
        ctx.save_for_backward(x, y, q)
        ctx.geo = (nx, ny, dx, dy)
        return tsc_deposit_raw(x, y, q, nx, ny, dx, dy)

    @staticmethod
    def backward(ctx, gplane):
        x, y, q = ctx.saved_tensors
        v, gx, gy = tsc_gather_raw(x, y, gplane.contiguous(), *ctx.geo, grad=True)
        return q * gx, q * gy, v, None, None, None, None


class TscGather(torch.autograd.Function):
    """v_p = sum_k W(x_p, y_p; k) plane[k]"""
    @staticmethod
    def forward(ctx, x, y, plane, nx, ny, dx, dy):
        x, y, plane = _f64(x, "x"), _f64(y, "y"), _f64(plane, "plane")
        ctx.save_for_backward(x, y, plane)
        ctx.geo = (nx, ny, dx, dy)
        return tsc_gather_raw(x, y, plane, nx, ny, dx, dy)

    @staticmethod
    def backward(ctx, gv):
        x, y, plane = ctx.saved_tensors
        gv = gv.contiguous()
        _, gx, gy = tsc_gather_raw(x, y, plane, *ctx.geo, grad=True)
        return gv * gx, gv * gy, tsc_deposit_raw(x, y, gv, *ctx.geo), None, None, None, None


def _boris_args(u, E, B, qm, dt):
    a = BorisArgs()
    for n, t in zip(("ux", "uy", "uz", "Ex", "Ey", "Ez", "Bx", "By", "Bz", "qm", "dt"), (*u, *E, *B, qm, dt)):
        setattr(a, n, t.data_ptr())
    a.np = u[0].numel()
    return a


class BorisFunction(torch.autograd.Function):
    """u' = boris(u, E, B, qm, dt) per particle: u, E, B as three 1-D tensors each"""
    @staticmethod
    def forward(ctx, ux, uy, uz, Ex, Ey, Ez, Bx, By, Bz, qm, dt):
        ins = [_f64(t, "boris input") for t in (ux, uy, uz, Ex, Ey, Ez, Bx, By, Bz, qm, dt)]
        ctx.save_for_backward(*ins)
        out = [t.clone() for t in ins[:3]]
        a = _boris_args(out, ins[3:6], ins[6:9], ins[9], ins[10])
        _call(g3d_amd.lib().g3d_boris, C.byref(a))
        return tuple(out)

    @staticmethod
    def backward(ctx, gx, gy, gz):
        ins = ctx.saved_tensors
        a = _boris_args(list(ins[:3]), ins[3:6], ins[6:9], ins[9], ins[10])
        ub = [(g if g is not None else torch.zeros_like(ins[0])).contiguous() for g in (gx, gy, gz)]
        outs = [torch.empty_like(ins[0]) for _ in range(11)]
        g = BorisAdj()
        for n, t in zip(("ubx", "uby", "ubz"), ub):
            setattr(g, n, t.data_ptr())
        for n, t in zip(("gux", "guy", "guz", "gEx", "gEy", "gEz", "gBx", "gBy", "gBz", "gqm", "gdt"), outs):
            setattr(g, n, t.data_ptr())
        _call(g3d_amd.lib().g3d_boris_vjp, C.byref(a), C.byref(g))
        return tuple(outs)
