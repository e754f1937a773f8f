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

WitnessEnergy: the mean p_z of the live bunch after a whole run as a function of the grid laser's envelope values
and the bunch's current-profile values, its gradient from the reverse of the march (g3d_sim_gradient).

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


# ---------------------------------------------------------------------------------------------- streaming push
_STATE = ("x", "y", "zeta", "z", "px", "py", "pz")


class SpushCtx(C.Structure):
    """g3d_spush_ctx (include/g3d.h)"""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("dx", C.c_double), ("dy", C.c_double), ("zeta0", C.c_double),
                ("ihz", C.c_double), ("n", C.c_int), ("zn", C.c_double), ("zn1", C.c_double),
                ("Delta", C.c_double), ("s1", C.c_double), ("sTrap", C.c_double), ("vz_beam", C.c_double),
                ("ia", C.c_double), ("F0", C.c_void_p), ("F1", C.c_void_p)] + \
               [(n, C.c_void_p) for n in _STATE] + \
               [("q", C.c_void_p), ("qm", C.c_void_p), ("w", C.c_void_p), ("alive", C.c_void_p), ("J", C.c_void_p),
                ("np", C.c_long)]


class SpushAdj(C.Structure):
    """g3d_spush_adj (include/g3d.h)"""
    _fields_ = [("b" + n, C.c_void_p) for n in _STATE] + [("bJ", C.c_void_p)] + \
               [("g" + n, C.c_void_p) for n in _STATE] + [("gF0", C.c_void_p), ("gF1", C.c_void_p)]


class HalfCell:
    """The half-cell and grid of a streaming push: a periodic nx x ny grid of steps dx, dy; slices n and n + 1 of a
    uniform ζ axis (slice k at zeta0 + k h); the step length Delta, its end s1, the trapping limit sTrap, the beam
    threshold vz_beam and the inverse cell-area factor ia of the deposit (src/sim_stream.hpp:20-29)"""
    def __init__(self, nx, ny, dx, dy, zeta0, h, n, Delta, s1, sTrap, vz_beam, ia):
        self.nx, self.ny, self.dx, self.dy = nx, ny, dx, dy
        self.zeta0, self.h, self.n = zeta0, h, n
        self.Delta, self.s1, self.sTrap, self.vz_beam, self.ia = Delta, s1, sTrap, vz_beam, ia

    def ctx(self, F0, F1, state, q, qm, w, alive, J):
        c = SpushCtx()
        c.nx, c.ny, c.dx, c.dy = self.nx, self.ny, self.dx, self.dy
        c.zeta0, c.ihz, c.n = self.zeta0, 1.0 / self.h, self.n
        c.zn, c.zn1 = self.zeta0 + self.n * self.h, self.zeta0 + (self.n + 1) * self.h
        c.Delta, c.s1, c.sTrap, c.vz_beam, c.ia = self.Delta, self.s1, self.sTrap, self.vz_beam, self.ia
        c.F0, c.F1 = F0.data_ptr(), F1.data_ptr()
        for nm, t in zip(_STATE, state):
            setattr(c, nm, t.data_ptr())
        c.q, c.qm, c.w, c.alive = q.data_ptr(), qm.data_ptr(), w.data_ptr(), alive.data_ptr()
        c.J = J.data_ptr() if J is not None else None
        c.np = state[0].numel()
        return c


def stream_push_raw(hc, F0, F1, state, q, qm, w, alive):
    """the pushed state (new tensors), the alive codes after the push and the deposit J [5][ny nx]"""
    out = [t.clone() for t in state]
    al = alive.clone()
    J = torch.zeros(5 * hc.nx * hc.ny, dtype=torch.float64, device=F0.device)
    c = hc.ctx(F0, F1, out, q, qm, w, al, J)
    _call(g3d_amd.lib().g3d_stream_push, C.byref(c))
    return out, al, J


class StreamPush(torch.autograd.Function):
    """(x, y, ζ, z, p_x, p_y, p_z) of the electrons after half-cell n, and their deposit J [5 ny nx], as functions of
    the state before it and of the fields F0, F1 [6 ny nx] of slices n and n + 1 (alive codes, q, q/m and w held
    fixed)"""
    @staticmethod
    def forward(ctx, hc, F0, F1, x, y, zeta, z, px, py, pz, q, qm, w, alive):
        state = [_f64(t, "state") for t in (x, y, zeta, z, px, py, pz)]
        F0, F1 = _f64(F0, "F0"), _f64(F1, "F1")
        q, qm, w = _f64(q, "q"), _f64(qm, "qm"), _f64(w, "w")
        ctx.hc = hc
        ctx.alive = alive.contiguous()
        ctx.save_for_backward(F0, F1, q, qm, w, *state)
        out, _, J = stream_push_raw(hc, F0, F1, state, q, qm, w, ctx.alive)
        return (*out, J)

    @staticmethod
    def backward(ctx, *grads):
        F0, F1, q, qm, w, *state = ctx.saved_tensors
        hc = ctx.hc
        zeros = torch.zeros_like(state[0])
        bs = [(g if g is not None else zeros).contiguous() for g in grads[:7]]
        bJ = (grads[7] if grads[7] is not None else torch.zeros(5 * hc.nx * hc.ny, dtype=torch.float64,
                                                                 device=F0.device)).contiguous()
        gs = [torch.empty_like(state[0]) for _ in _STATE]
        gF0, gF1 = torch.zeros_like(F0), torch.zeros_like(F1)
        c = hc.ctx(F0, F1, list(state), q, qm, w, ctx.alive, None)
        g = SpushAdj()
        for nm, t in zip(_STATE, bs):
            setattr(g, "b" + nm, t.data_ptr())
        g.bJ = bJ.data_ptr()
        for nm, t in zip(_STATE, gs):
            setattr(g, "g" + nm, t.data_ptr())
        g.gF0, g.gF1 = gF0.data_ptr(), gF1.data_ptr()
        _call(g3d_amd.lib().g3d_stream_push_vjp, C.byref(c), C.byref(g))
        return (None, gF0, gF1, *gs, None, None, None, None)


# ---------------------------------------------------------------------------------------------- RK4 bunch push
_BSTATE = ("x", "y", "zeta", "px", "py", "pz")


class BpushCtx(C.Structure):
    """g3d_bpush_ctx (include/g3d.h)"""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nsl", C.c_int)] + \
               [(n, C.c_double) for n in ("dx", "dy", "zeta0", "ihz")] + [("F", C.c_void_p)] + \
               [("laser_on", C.c_int), ("laser_pol", C.c_int)] + \
               [(n, C.c_double) for n in ("a0", "kL", "tau", "cep", "eta_c", "kx", "kz", "X0", "zc", "s0", "s1",
                                          "dz_push", "push_phase", "hmin", "zhand", "zmax", "ylo", "yhi", "vz_min")] + \
               [(n, C.c_void_p) for n in _BSTATE + ("z", "rem", "qm", "alive")] + [("np", C.c_long)]


class BpushAdj(C.Structure):
    """g3d_bpush_adj (include/g3d.h)"""
    _fields_ = [("b" + n, C.c_void_p) for n in _BSTATE] + [("g" + n, C.c_void_p) for n in _BSTATE] + \
               [("gF", C.c_void_p)]


class PushGeometry:
    """The grid, ζ axis, laser and step of an RK4 bunch push: a periodic nx x ny grid of steps dx, dy; nsl slices
    at zeta0 + k h; the plane wave (laser_on, laser_pol, a0, kL, tau, cep, eta_c, kx, kz, as g3d_config) with the
    window offset X0 and centre zc; the step s0 -> s1 with dz_push, push_phase, hmin, zhand, zmax, ylo, yhi, vz_min
    (src/sim_kernels.hpp:10-17)"""
    def __init__(self, **kw):
        self.kw = kw

    def ctx(self, F, state, z, rem, qm, alive):
        c = BpushCtx()
        kw = dict(self.kw)
        c.ihz = 1.0 / kw.pop("h")
        for k, v in kw.items():
            setattr(c, k, v)
        c.F = F.data_ptr()
        for nm, t in zip(_BSTATE, state):
            setattr(c, nm, t.data_ptr())
        c.z, c.rem, c.qm, c.alive = z.data_ptr(), rem.data_ptr(), qm.data_ptr(), alive.data_ptr()
        c.np = state[0].numel()
        return c


def bunch_push_raw(geo, F, state, z, rem, qm, alive):
    """the pushed state (new tensors) and the new z, rem, alive"""
    out = [t.clone() for t in state]
    z2, rem2, al = z.clone(), rem.clone(), alive.clone()
    c = geo.ctx(F, out, z2, rem2, qm, al)
    _call(g3d_amd.lib().g3d_bunch_push_piece, C.byref(c))
    return out, z2, rem2, al


class BunchPush(torch.autograd.Function):
    """(x, y, ζ, p_x, p_y, p_z) of beam particles after the RK4 push of a step, as functions of the state before it
    and of the field planes F [6 nsl ny nx] (z, rem, q/m and the alive codes held fixed)"""
    @staticmethod
    def forward(ctx, geo, F, x, y, zeta, px, py, pz, z, rem, qm, alive):
        state = [_f64(t, "state") for t in (x, y, zeta, px, py, pz)]
        F = _f64(F, "F")
        ctx.geo = geo
        ctx.fixed = (z.contiguous(), rem.contiguous(), qm.contiguous(), alive.contiguous())
        ctx.save_for_backward(F, *state)
        out, _, _, _ = bunch_push_raw(geo, F, state, *ctx.fixed)
        return tuple(out)

    @staticmethod
    def backward(ctx, *grads):
        F, *state = ctx.saved_tensors
        zeros = torch.zeros_like(state[0])
        bs = [(g if g is not None else zeros).contiguous() for g in grads]
        gF = torch.zeros_like(F)
        gs = bunch_push_vjp_raw(ctx.geo, F, state, *ctx.fixed, bs, gF)
        return (None, gF, *gs, None, None, None, None)


def bunch_push_vjp_raw(geo, F, state, z, rem, qm, alive, bs, gF):
    """the adjoints of the state before the push (new tensors) from those of the pushed state bs; the adjoint of F
    is added to gF.  state, z, rem and alive are the ones before the push (left unchanged)"""
    gs = [torch.empty_like(state[0]) for _ in _BSTATE]
    c = geo.ctx(F, [t.clone() for t in state], z.clone(), rem.clone(), qm, alive.clone())
    g = BpushAdj()
    for nm, t in zip(_BSTATE, bs):
        setattr(g, "b" + nm, t.data_ptr())
    for nm, t in zip(_BSTATE, gs):
        setattr(g, "g" + nm, t.data_ptr())
    g.gF = gF.data_ptr()
    _call(g3d_amd.lib().g3d_bunch_push_vjp, C.byref(c), C.byref(g))
    return gs


class BunchMarch(torch.autograd.Function):
    """(x, y, ζ, p_x, p_y, p_z) of beam particles after len(geos) RK4 pushes, one PushGeometry per s-step, in the
    field planes F [6 nsl ny nx] held over the steps.  The forward keeps the particle state only at every
    `every`-th step; the reverse sweep recomputes each segment of steps from its checkpoint, then runs the steps'
    reverse pushes rear to front (memory: one state per checkpoint plus one segment).  z, rem and the alive codes
    move with the steps and are not differentiated."""
    @staticmethod
    def forward(ctx, geos, every, F, x, y, zeta, px, py, pz, z, rem, qm, alive):
        if every < 1:
            raise ValueError("BunchMarch: every must be >= 1")
        state = [_f64(t, "state") for t in (x, y, zeta, px, py, pz)]
        F = _f64(F, "F")
        fix = (z.contiguous(), rem.contiguous(), alive.contiguous())
        qm = qm.contiguous()
        ckpts = []
        for t, geo in enumerate(geos):
            if t % every == 0:
                ckpts.append((t, [u.clone() for u in state], *[u.clone() for u in fix]))
            state, *fix = _push_step(geo, F, state, fix, qm)
        ctx.geos, ctx.every, ctx.ckpts, ctx.np = list(geos), every, ckpts, state[0].numel()
        ctx.save_for_backward(F, qm)
        return tuple(state)

    @staticmethod
    def backward(ctx, *grads):
        if ctx.ckpts is None:
            raise RuntimeError("BunchMarch: a second backward through the same graph is not supported (the particle "
                               "checkpoints are released by the first); run the forward again")
        F, qm = ctx.saved_tensors
        geos, every = ctx.geos, ctx.every
        zeros = torch.zeros(ctx.np, dtype=torch.float64, device=F.device)
        bs = [(g if g is not None else zeros).contiguous() for g in grads]
        gF = torch.zeros_like(F)
        for t0, state, z, rem, alive in reversed(ctx.ckpts):
            seg = []                                   # the inputs of each step of the segment
            fix = (z, rem, alive)
            for t in range(t0, min(t0 + every, len(geos))):
                seg.append((state, fix))
                state, *fix = _push_step(geos[t], F, state, fix, qm)
            for k in range(len(seg) - 1, -1, -1):
                st, (zz, rr, al) = seg[k]
                bs = bunch_push_vjp_raw(geos[t0 + k], F, st, zz, rr, qm, al, bs, gF)
        ctx.ckpts = None
        return (None, None, gF, *bs, None, None, None, None)


def _push_step(geo, F, state, fix, qm):
    z, rem, alive = fix
    out, z2, rem2, al = bunch_push_raw(geo, F, state, z, rem, qm, alive)
    return out, z2, rem2, al


class WitnessEnergy(torch.autograd.Function):
    """The mean p_z of the live bunch after the march of a deck, as a function of the envelope values of its first
    grid laser and the values of its bunch's current profile (float64 tensors on any device).  make_deck(env, prof)
    builds the deck from numpy arrays; steps and every as g3d_sim_gradient.  The forward runs the whole march and
    its reverse on the GPU (Sim.gradient: the transposed field updates, the reverse pushes, a state checkpoint
    every `every` steps with each segment run again); backward scales the stored gradient by the incoming one."""
    @staticmethod
    def forward(ctx, env, prof, make_deck, steps, every):
        deck = make_deck(env.detach().cpu().numpy().astype("float64"), prof.detach().cpu().numpy().astype("float64"))
        sim = g3d_amd.Sim(deck)
        try:
            loss, genv, gprof = sim.gradient(int(steps), int(every))
        finally:
            sim.close()
        ctx.save_for_backward(torch.from_numpy(genv).to(env.device), torch.from_numpy(gprof).to(prof.device))
        return torch.tensor(loss, dtype=torch.float64, device=env.device)

    @staticmethod
    def backward(ctx, g):
        genv, gprof = ctx.saved_tensors
        return g * genv, g * gprof, None, None, None
