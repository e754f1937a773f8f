"""The adjoint of the march (g3d_march_adjoint) behind torch.autograd (g3d_autograd.MarchFunction).

- the dot-product identity <M x, y> = <x, M^T y> for random x and y (FP64, to rounding);
- the gradient of a scalar loss against central finite differences of the CPU oracle's march (the restatement
  pinned bit for bit to the reference's build), within 1e-6 relative, in random directions;
- the tangent-linear run: the march is linear in its inputs, so M(x + dx) - M(x) = M(dx) to rounding."""
import numpy as np
import pytest
import torch

import g3d_amd
import g3d_autograd as ga
import golden_inputs as gi
import oraclepy as orc

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not orc.available(), reason="oracle not built")]
FIELDS = ("Rx", "Ry", "Ez", "Bz")
CURRENTS = ("j0x", "j0y", "jkx", "jky", "jkz")


def setup(seed=5):
    deck = gi.small_deck()
    sim = g3d_amd.Sim(deck)
    geom = ga.MarchGeometry(sim, gi.MARCH_D)
    r = np.random.default_rng(seed)
    ns, nk = geom.ns, geom.nk

    def rnd(rows):
        return r.standard_normal((rows, nk)) + 1j * r.standard_normal((rows, nk))
    x = {n: rnd(ns) for n in FIELDS}
    x.update({n: rnd(ns + 1) for n in CURRENTS})
    return deck, geom, x, rnd


def gpu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def run(geom, x):
    out = ga.march(geom, [gpu(x[n]) for n in FIELDS], [gpu(x[n]) for n in CURRENTS])
    return [o.cpu().numpy() for o in out]


def inner(u, v):
    return sum(float(np.real(np.vdot(a, b))) for a, b in zip(u, v))


def test_march_adjoint_dot_product():
    deck, geom, x, rnd = setup(5)
    y = [rnd(geom.ns) for _ in range(6)]
    mx = run(geom, x)
    xs = [gpu(x[n]).requires_grad_() for n in FIELDS + CURRENTS]
    out = ga.march(geom, xs[:4], xs[4:])
    loss = sum(torch.sum(torch.real(torch.conj(gpu(w)) * o)) for w, o in zip(y, out))
    loss.backward()
    lhs = inner(mx, y)
    rhs = inner([x[n] for n in FIELDS + CURRENTS], [t.grad.cpu().numpy() for t in xs])
    print("dot product", lhs, rhs)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)
    assert abs(lhs - float(loss.item())) <= 1e-12 * max(abs(lhs), 1.0)


def test_march_gradient_against_oracle_fd():
    """loss = Re sum conj(w) M(x) over the six output planes; the adjoint gradient against central differences
    (step 1e-6 relative to the input scale) of the oracle's march, in four random directions"""
    deck, geom, x, rnd = setup(7)
    w = [rnd(geom.ns) for _ in range(6)]
    g = orc.Grid(deck)
    names = ("Rx", "Ry", "Ez", "Bz", "Lx", "Ly")

    def loss_oracle(z):
        arr = dict(z)
        arr["Lx"] = np.zeros_like(z["Rx"])
        arr["Ly"] = np.zeros_like(z["Rx"])
        out = orc.march(g, {k: v.ravel() for k, v in arr.items()}, gi.MARCH_D)
        return sum(float(np.real(np.vdot(wi, out[n].reshape(wi.shape)))) for wi, n in zip(w, names))

    xs = [gpu(x[n]).requires_grad_() for n in FIELDS + CURRENTS]
    out = ga.march(geom, xs[:4], xs[4:])
    loss = sum(torch.sum(torch.real(torch.conj(gpu(wi)) * o)) for wi, o in zip(w, out))
    loss.backward()
    grad = {n: t.grad.cpu().numpy() for n, t in zip(FIELDS + CURRENTS, xs)}
    r = np.random.default_rng(11)
    eps = 1e-6
    for _ in range(4):
        v = {n: r.standard_normal(a.shape) + 1j * r.standard_normal(a.shape) for n, a in x.items()}
        fd = (loss_oracle({n: x[n] + eps * v[n] for n in x}) - loss_oracle({n: x[n] - eps * v[n] for n in x})) / (2 * eps)
        ad = sum(float(np.real(np.vdot(grad[n], v[n]))) for n in x)
        print("directional derivative: adjoint %.12e  finite difference %.12e" % (ad, fd))
        assert abs(ad - fd) <= 1e-6 * abs(fd)


def test_march_tangent_linear():
    deck, geom, x, rnd = setup(9)
    dx = {n: 1e-3 * (rnd(a.shape[0])) for n, a in x.items()}
    a = run(geom, x)
    b = run(geom, {n: x[n] + dx[n] for n in x})
    t = run(geom, dx)
    for p, q, s in zip(a, b, t):
        assert np.linalg.norm((q - p) - s) <= 1e-9 * np.linalg.norm(s)


def test_march_adjoint_front_slice():
    """E_z of the front slice is zero whatever the inputs (L = E_z = B_z = 0 there): its gradient is zero"""
    deck, geom, x, rnd = setup(3)
    ns, nk = geom.ns, geom.nk
    xs = [gpu(x[n]).requires_grad_() for n in FIELDS + CURRENTS]
    out = ga.march(geom, xs[:4], xs[4:])
    torch.sum(torch.real(out[2][ns - 1])).backward()
    assert all(float(torch.abs(t.grad).max()) == 0.0 for t in xs)


# ------------------------------------------------------------------------------ the particle-grid pieces
NX, NY, DX, DY = 8, 8, 0.7, 0.9


def particles(r, n):
    x = torch.from_numpy(r.uniform(0, NX * DX, n)).cuda()
    y = torch.from_numpy(r.uniform(0, NY * DY, n)).cuda()
    return x, y


def test_tsc_pair_dot_product():
    """<deposit(q), P> = <q, gather(P)>: the gather is the deposit's transpose"""
    r = np.random.default_rng(21)
    x, y = particles(r, 300)
    q = torch.from_numpy(r.standard_normal(300)).cuda()
    P = torch.from_numpy(r.standard_normal(NX * NY)).cuda()
    lhs = float(torch.dot(ga.tsc_deposit_raw(x, y, q, NX, NY, DX, DY), P))
    rhs = float(torch.dot(q, ga.tsc_gather_raw(x, y, P, NX, NY, DX, DY)))
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


def test_tsc_deposit_weights():
    """the deposit of unit charges sums to the particle count, and one particle's weights are the closed-form TSC
    weights (src/sim_kernels.hpp:21-31; periodic, nodes at i dx)"""
    r = np.random.default_rng(22)
    x, y = particles(r, 50)
    plane = ga.tsc_deposit_raw(x, y, torch.ones(50, dtype=torch.float64, device="cuda"), NX, NY, DX, DY)
    assert abs(float(plane.sum()) - 50) < 1e-12
    xi, yi = 2.3 * DX, 5.9 * DY            # one particle: weights 1/2 (1/2 - f)^2, 3/4 - f^2, 1/2 (1/2 + f)^2
    one = ga.tsc_deposit_raw(torch.tensor([xi], dtype=torch.float64, device="cuda"),
                             torch.tensor([yi], dtype=torch.float64, device="cuda"),
                             torch.ones(1, dtype=torch.float64, device="cuda"), NX, NY, DX, DY).cpu().numpy()
    def w(s):
        i0 = int(np.floor(s + 0.5)); f = s - i0
        return i0, [0.5 * (0.5 - f) ** 2, 0.75 - f * f, 0.5 * (0.5 + f) ** 2]
    ix, wx = w(xi / DX); jy, wy = w(yi / DY)
    want = np.zeros(NX * NY)
    for v in range(3):
        for u in range(3):
            want[((jy - 1 + v) % NY) * NX + (ix - 1 + u) % NX] += wx[u] * wy[v]
    np.testing.assert_allclose(one, want, rtol=0, atol=1e-15)


def test_tsc_gradcheck():
    """deposit and gather differentiated in the particle positions, the values and the plane: the hand-written
    backward against torch's central differences"""
    r = np.random.default_rng(23)
    x, y = particles(r, 6)
    q = torch.from_numpy(r.standard_normal(6)).cuda()
    P = torch.from_numpy(r.standard_normal(NX * NY)).cuda()
    for t in (x, y, q, P):
        t.requires_grad_()
    assert torch.autograd.gradcheck(lambda a, b, c: ga.TscDeposit.apply(a, b, c, NX, NY, DX, DY), (x, y, q),
                                    eps=1e-6, atol=1e-8, rtol=1e-6)
    assert torch.autograd.gradcheck(lambda a, b, c: ga.TscGather.apply(a, b, c, NX, NY, DX, DY), (x, y, P),
                                    eps=1e-6, atol=1e-8, rtol=1e-6)


def test_boris_vjp_gradcheck():
    """the reverse Boris rotation against central differences of the forward one (every input: u, E, B, qm, dt),
    and the forward against the oracle's boris bit for bit"""
    r = np.random.default_rng(24)
    n = 5
    ins = [torch.from_numpy(r.standard_normal(n) * s).cuda().requires_grad_()
           for s in (2.0, 2.0, 2.0, 0.5, 0.5, 0.5, 0.3, 0.3, 0.3)]
    qm = torch.from_numpy(-np.ones(n) + 0.1 * r.standard_normal(n)).cuda().requires_grad_()
    dt = torch.from_numpy(r.uniform(0.1, 0.5, n)).cuda().requires_grad_()
    assert torch.autograd.gradcheck(ga.BorisFunction.apply, (*ins, qm, dt), eps=1e-6, atol=1e-8, rtol=1e-6)
    import ctypes as C
    out = ga.BorisFunction.apply(*ins, qm, dt)
    for p in range(n):
        u = np.array([float(t[p]) for t in ins[:3]])
        E = np.array([float(t[p]) for t in ins[3:6]]); B = np.array([float(t[p]) for t in ins[6:9]])
        orc.lib().or_boris(u.ctypes.data_as(C.c_void_p), E.ctypes.data_as(C.c_void_p), B.ctypes.data_as(C.c_void_p),
                           float(qm[p]), float(dt[p]))
        assert [float(o[p]) for o in out] == list(u)


# ------------------------------------------------------------------------------ the streaming push
def spush_setup():
    """the golden stream input (64 electrons of half-cell STREAM_N of the 8 x 8 x 12 deck, random fields) as
    FP64 planes of slices n, n + 1 and the state tensors"""
    deck = gi.small_deck()
    sim = g3d_amd.Sim(deck)
    sim.init()
    F, P = gi.stream_input(sim)
    info = sim.info()
    ns, nk = info["ns"], info["nk"]
    n = gi.STREAM_N
    Fv = F.reshape(6, ns + 2, nk).astype(np.float64)
    g = orc.Grid(deck)
    c = deck.c
    hc = ga.HalfCell(g.nx, g.ny, g.dx, g.dy, 0.0, 1 / g.ih, n, 5.0, 5.0, 5.0 + (c.trap_advance - 1) * 5.0, c.vz_beam,
                     g.iA[0])
    return deck, sim, F, P, hc, gpu(Fv[:, n].ravel()), gpu(Fv[:, n + 1].ravel()), info


def test_stream_push_against_oracle():
    """the forward streaming push (g3d_stream_push) against the oracle's stream pass on the golden input: the state
    and alive codes of every electron (waiting, entering, trapped, dead), the deposit to 1e-13 of its largest value"""
    deck, sim, F, P, hc, F0, F1, info = spush_setup()
    want, J_want, _, _ = orc.stream_halfcell(orc.Grid(deck), deck, F, P, gi.STREAM_N, 0.0, 5.0, info["X0"], info["Y0"])
    state = [gpu(P[f].astype(np.float64)) for f in ("x", "y", "zeta", "z", "px", "py", "pz")]
    out, alive, J = ga.stream_push_raw(hc, F0, F1, state, gpu(P["q"]), gpu(P["qm"]), gpu(P["w"]),
                                       gpu(P["alive"].astype(np.int32)))
    for f, t in zip(("x", "y", "zeta", "z", "px", "py", "pz"), out):
        got = t.cpu().numpy()
        assert np.allclose(got, want[f], rtol=1e-14, atol=1e-14), (f, np.abs(got - want[f]).max())
    assert np.array_equal(alive.cpu().numpy(), want["alive"])
    J = J.cpu().numpy()
    assert np.abs(J - J_want).max() <= 1e-13 * np.abs(J_want).max()


def test_stream_push_gradcheck():
    """the reverse streaming push against central differences of the forward one, in every input (the state and
    both field planes): electrons inside the half-cell (ζ between the slices), a fast one that is trapped in the
    push and deposits its beam current, slow ones with a strong transverse kick"""
    deck, sim, F, P, hc, F0, F1, info = spush_setup()
    r = np.random.default_rng(31)
    m = 5
    zn, h = hc.zeta0 + hc.n * hc.h, hc.h
    x = gpu(r.uniform(0, hc.nx * hc.dx, m)); y = gpu(r.uniform(0, hc.ny * hc.dy, m))
    zeta = gpu(zn + h * r.uniform(0.2, 0.8, m)); z = gpu(r.uniform(0.0, 1.0, m))
    px = gpu(0.3 * r.standard_normal(m)); py = gpu(0.3 * r.standard_normal(m)); pz = gpu(0.2 * r.standard_normal(m))
    pz[0] = 30.0
    ones = torch.ones(m, dtype=torch.float64, device="cuda")
    q, qm, w = -ones, -ones, 1e-2 * ones
    alive = torch.ones(m, dtype=torch.int32, device="cuda")
    ins = [F0.clone().requires_grad_(), F1.clone().requires_grad_()] + \
          [t.clone().requires_grad_() for t in (x, y, zeta, z, px, py, pz)]
    out, al, _ = ga.stream_push_raw(hc, ins[0].detach(), ins[1].detach(), [t.detach() for t in ins[2:]], q, qm, w,
                                    alive)
    assert int(al[0]) == 5 and (al[1:] == 1).all()          # the fast one is trapped, the others stream
    f = lambda *a: ga.StreamPush.apply(hc, *a, q, qm, w, alive)
    assert torch.autograd.gradcheck(f, tuple(ins), eps=1e-6, atol=1e-8, rtol=1e-5)


# ------------------------------------------------------------------------------ the RK4 bunch push
def bpush_geometry(deck, sim, s0, s1):
    g = orc.Grid(deck)
    c = deck.c
    L = orc.laser_of(deck)
    info = sim.info()
    return ga.PushGeometry(nx=g.nx, ny=g.ny, nsl=info["ns"] + 2, dx=g.dx, dy=g.dy, zeta0=0.0, h=1 / g.ih,
                           laser_on=L.on, laser_pol=L.pol, a0=L.a0, kL=L.kL, tau=L.tau, cep=L.cep, eta_c=L.eta_c,
                           kx=L.kx, kz=L.kz, X0=info["X0"], zc=c.zeta_center, s0=s0, s1=s1,
                           dz_push=c.dz_push or c.dz, push_phase=c.push_phase,
                           hmin=float(np.min(np.diff(g.zeta))), zhand=-1e300, zmax=float(g.zeta[-1]), ylo=-1e300,
                           yhi=1e300, vz_min=0.0)


def test_bunch_push_piece_against_oracle():
    """the forward RK4 piece (g3d_bunch_push_piece) against the oracle's bunch push on the golden input (32 beam
    particles with the oblique plane wave, one slow with many sub-steps); the laser's exp, cos and sin may differ
    from glibc's in the last bit"""
    deck = gi.bunch_small_deck()
    sim = g3d_amd.Sim(deck)
    sim.init()
    F, P = gi.push_input(sim)
    geo = bpush_geometry(deck, sim, 0.0, 20.0)
    info = sim.info()
    want = orc.bunch_push(orc.Grid(deck), deck, F, P, 0.0, 20.0, info["X0"], info["Y0"])
    state = [gpu(P[f].astype(np.float64)) for f in ("x", "y", "zeta", "px", "py", "pz")]
    out, z, rem, alive = ga.bunch_push_raw(geo, gpu(F.astype(np.float64)), state, gpu(P["z"]), gpu(P["rem"]),
                                           gpu(P["qm"]), gpu(P["alive"].astype(np.int32)))
    for f, t in zip(("x", "y", "zeta", "px", "py", "pz"), out):
        assert np.allclose(t.cpu().numpy(), want[f], rtol=1e-13, atol=0), f
    assert np.array_equal(alive.cpu().numpy(), want["alive"])
    assert np.allclose(z.cpu().numpy(), want["z"], rtol=0, atol=0)


def test_bunch_push_gradcheck():
    """the reverse RK4 push against central differences of the forward one, in the state and every field value:
    four particles from p_z = 3 (many laser-limited sub-steps) to 1000 (one), in the oblique plane wave.  atol
    1e-7: the central difference of an output near 1000 with step 1e-6 carries about 1e-7 of rounding"""
    deck = gi.bunch_small_deck()
    sim = g3d_amd.Sim(deck)
    sim.init()
    info = sim.info()
    geo = bpush_geometry(deck, sim, 0.0, 20.0)
    r = np.random.default_rng(41)
    nsl, nk = info["ns"] + 2, info["nk"]
    F = gpu(0.01 * r.standard_normal(6 * nsl * nk)).requires_grad_()
    m = 4
    x = gpu(8 + r.standard_normal(m)); y = gpu(8 + r.standard_normal(m))
    zeta = gpu(3.75 + 0.5 * r.standard_normal(m))
    px = gpu(1e-2 * r.standard_normal(m)); py = gpu(1e-2 * r.standard_normal(m))
    pz = gpu(np.array([3.0, 10.0, 50.0, 1000.0]))
    fixed = (torch.zeros(m, dtype=torch.float64, device="cuda"), torch.zeros(m, dtype=torch.float64, device="cuda"),
             -torch.ones(m, dtype=torch.float64, device="cuda"), torch.ones(m, dtype=torch.int32, device="cuda"))
    ins = (F,) + tuple(t.requires_grad_() for t in (x, y, zeta, px, py, pz))
    out, _, _, al = ga.bunch_push_raw(geo, F.detach(), [t.detach() for t in ins[1:]], *fixed)
    assert (al == 1).all()
    f = lambda *a: ga.BunchPush.apply(geo, *a, *fixed)
    assert torch.autograd.gradcheck(f, ins, eps=1e-6, atol=1e-7, rtol=1e-5)


def test_bunch_march_checkpoints():
    """six RK4 pushes of the 32 golden beam particles (the oblique plane wave, one slow particle with many
    sub-steps) with particle checkpoints every k steps (ga.BunchMarch): the gradient of a random linear functional
    of the final state is the same for every k and equals the step-by-step autograd chain; in the initial state it
    matches central differences (step 1e-6) of the oracle's chained bunch push within 1e-6 relative, and in the
    field planes central differences of the forward chain"""
    deck = gi.bunch_small_deck()
    sim = g3d_amd.Sim(deck)
    sim.init()
    info = sim.info()
    F32, P = gi.push_input(sim)
    S, ds = 6, 20.0
    geos = [bpush_geometry(deck, sim, ds * t, ds * (t + 1)) for t in range(S)]
    names = ("x", "y", "zeta", "px", "py", "pz")
    r = np.random.default_rng(77)
    cot = [r.standard_normal(len(P)) for _ in names]
    F = gpu(F32.astype(np.float64))
    fixed = (gpu(P["z"]), gpu(P["rem"]), gpu(P["qm"]), gpu(P["alive"].astype(np.int32)))
    z, rem, qm, alive = fixed

    def grads(every):
        ins = [F.clone().requires_grad_()] + [gpu(P[n].astype(np.float64)).requires_grad_() for n in names]
        if every is None:                             # one autograd node per step
            st, zz, rr, al = ins[1:], z, rem, alive
            for geo in geos:
                st2 = ga.BunchPush.apply(geo, ins[0], *st, zz, rr, qm, al)
                _, zz, rr, al = ga.bunch_push_raw(geo, ins[0].detach(), [t.detach() for t in st], zz, rr, qm, al)
                st = st2
        else:
            st = ga.BunchMarch.apply(geos, every, ins[0], *ins[1:], z, rem, qm, alive)
        loss = sum((t * gpu(c)).sum() for t, c in zip(st, cot))
        g = torch.autograd.grad(loss, ins)
        return float(loss), [t.cpu().numpy() for t in g]

    L0, g_chain = grads(None)
    for every in (1, 2, 4, S, S + 3):
        L, g = grads(every)
        assert L == L0
        for a, b in zip(g, g_chain):
            assert np.allclose(a, b, rtol=1e-12, atol=1e-15 * np.abs(b).max()), every

    def oracle_loss(Pin):
        p = Pin
        for t in range(S):
            p = orc.bunch_push(orc.Grid(deck), deck, F32, p, ds * t, ds * (t + 1), info["X0"], info["Y0"])
        assert (p["alive"] == 1).all()
        return sum(float(np.dot(p[n], c)) for n, c in zip(names, cot))

    d = [r.standard_normal(len(P)) for _ in names]
    eps = 1e-6
    Pp, Pm = P.copy(), P.copy()
    for n, v in zip(names, d):
        Pp[n] += eps * v
        Pm[n] -= eps * v
    fd = (oracle_loss(Pp) - oracle_loss(Pm)) / (2 * eps)
    an = sum(float(np.dot(gs, v)) for gs, v in zip(g_chain[1:], d))
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)

    dF = r.standard_normal(F32.size)
    def gpu_loss(Fv):
        st, zz, rr, al = [gpu(P[n].astype(np.float64)) for n in names], z, rem, alive
        for geo in geos:
            st, zz, rr, al = ga.bunch_push_raw(geo, gpu(Fv), st, zz, rr, qm, al)
        return sum(float((t * gpu(c)).sum()) for t, c in zip(st, cot))
    F64 = F32.astype(np.float64)
    fdF = (gpu_loss(F64 + eps * dF) - gpu_loss(F64 - eps * dF)) / (2 * eps)
    anF = float(np.dot(g_chain[0], dF))
    assert abs(fdF - anF) <= 1e-6 * abs(anF), (fdF, anF)
