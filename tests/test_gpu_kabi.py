"""The kernel-level C ABI (include/g3d.h: g3d_stream_halfcell, g3d_stream_spectra, g3d_real_fields_slice,
g3d_bunch_push, g3d_bunch_currents, g3d_sponge_real_fields, g3d_comoving, g3d_bunch_centroid, g3d_window_shift) driven as the reference's own member
functions drive their device code, on caller-owned device buffers in the reference's layouts, against the
reference's CPU build (oracle/_ref) on one thread: the same inputs, the same outputs.  The geometry tables of the
constructor come from the engine (their values are held to the reference by the engine tests)."""
import ctypes as C
import math

import numpy as np
import pytest

import g3d_amd
import refpy
from decks import bunch_deck, particle_parity, plasma_deck

pytestmark = pytest.mark.gpu
D = g3d_amd.DevArray


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


class Tables:
    """device copies of the constructor's tables of an engine simulation"""

    def __init__(self, sim):
        self.sim = sim
        self.info = sim.info()
        self.d = {n: D(sim.array(n)) for n in ("kx", "ky", "kxs", "kys", "filt", "filty", "hcg", "wlo", "whi", "zs",
                                                "iA", "iD")}

    def spec(self, deck, plan, keep0=0):
        i = self.info
        return g3d_amd.SpecCtx(nx=deck.c.nx, ny=deck.c.ny, ns=i["ns"], N=i["N"], a=0, nk=i["nk"],
                               kx=self.d["kx"].ptr.value, ky=self.d["ky"].ptr.value, kxs=self.d["kxs"].ptr.value,
                               kys=self.d["kys"].ptr.value, filt=self.d["filt"].ptr.value,
                               filty=self.d["filty"].ptr.value, hcg=self.d["hcg"].ptr.value,
                               wlo=self.d["wlo"].ptr.value, whi=self.d["whi"].ptr.value,
                               a2=deck.c.smooth_length ** 2, keep0=keep0, plan=plan.h.value)

    def zax(self, deck):
        N = self.info["N"]
        return g3d_amd.Axis(X=self.d["zs"].ptr.value, n=N, uniform=1, x0=0.0, ih=1.0 / (deck.zeta[-1] / (N - 1)))

    def push(self, deck, F, s1, vz_min=0.0):
        i = self.info
        zeta = deck.zeta
        c = g3d_amd.PushCtx(F=F.ptr.value, a=0, ns=i["ns"], N=i["N"], nx=deck.c.nx, ny=deck.c.ny, self=1, yf=1,
                            nk=i["nk"], dx=i["dx"], dy=i["dy"], Lx=deck.c.Lx, Ly=deck.c.Ly)
        c.zax = self.zax(deck)
        c.yax = g3d_amd.Axis(X=None, n=deck.c.ny + 1, uniform=1, x0=0.0, ih=1.0 / i["dy"])
        c.X0, c.Y0, c.zc = i["X0"], i["Y0"], deck.c.zeta_center
        c.dz_push, c.push_phase, c.zhand, c.zmax = deck.c.dz_push or deck.c.dz, deck.c.push_phase, -1e300, zeta[-1]
        c.ylo, c.yhi, c.s1, c.vz_min, c.freeze = -1e300, 1e300, s1, vz_min, 0
        c.hmin = float(np.diff(zeta).min())
        return c


def test_halfcells_spectra_and_real_fields_against_reference():
    """twelve half-cells of the streaming plasma from the front (pass 0 in the fields of init), then the spectra of the
    last half-cell's current and the real fields of a slice"""
    mk = lambda: plasma_deck(nx=32, ny=32, nz=48, a0=1.0, n0=1e-2)
    deck = mk()
    refpy.set_threads(1)
    ref = refpy.RefSim(mk())
    ref.init()
    tabs = Tables(g3d_amd.Sim(mk()))
    info = ref.info()
    N, ns, nk = info["N"], info["ns"], info["nk"]
    s0, s1 = 0.0, deck.c.dz
    ref.call("load_plasma", s0, s1)
    strm = ref.array("strm")
    parts = g3d_amd.DeviceParts(strm)
    F = D(ref.array("F"))
    J = D(nbytes=5 * nk * 8)
    sc = g3d_amd.StreamCtx(pc=tabs.push(deck, F, s1), zs=tabs.d["zs"].ptr.value, J=J.ptr.value,
                           ia=tabs.d["iA"].ptr.value, id=tabs.d["iD"].ptr.value, Delta=s1 - s0, s1=s1,
                           sTrap=s1 + (deck.c.trap_advance - 1) * (s1 - s0), vz_beam=deck.c.vz_beam)
    for n in range(N - 2, N - 14, -1):
        ref.call("stream_pass", n, 0, s0, s1)
        sc.n = n
        g3d_amd.lib().g3d_memset0(J.ptr, 5 * nk * 8)
        g3d_amd.kernel_call("g3d_stream_halfcell", C.byref(sc), C.byref(parts.c), C.byref(parts.c), len(strm))
        jr, jg = ref.array("jsr"), J.get(np.float64)
        assert rel(jg, jr) <= 1e-12, (n, rel(jg, jr))
    got, want = parts.records(), ref.array("strm")
    np.testing.assert_array_equal(got["alive"], want["alive"])
    particle_parity(got, want, deck)
    assert np.abs(jr).max() > 0
    # the spectra of the last half-cell's current (the reference's stream_spectra reads its jsr)
    plan = g3d_amd.FftPlan(deck.c.nx, deck.c.ny)
    spec = tabs.spec(deck, plan, keep0=deck.c.box_mean)
    jsn = D(nbytes=3 * nk * 16)
    scr = D(nbytes=6 * nk * 16)
    ref.call("stream_spectra")
    g3d_amd.kernel_call("g3d_stream_spectra", C.byref(spec), J.ptr, jsn.ptr, scr.ptr)
    sr, sg = ref.array("jsn"), jsn.get(np.complex128)
    assert rel(sg, sr) <= 1e-12, rel(sg, sr)
    # real fields of slice n from the spectra of init
    n = N // 2
    arrs = {c: D(ref.array(c)) for c in ("Rx", "Ry", "Lx", "Ly", "Ez", "Bz")}
    off = lambda c: arrs[c].ptr.value + n * nk * 16
    Fg = D(nbytes=6 * (ns + 2) * nk * 4)
    ref.call("real_fields_slice", n)
    g3d_amd.kernel_call("g3d_real_fields_slice", C.byref(spec), *[C.c_void_p(off(c)) for c in
                                                                  ("Rx", "Ry", "Lx", "Ly", "Ez", "Bz")],
                        C.c_void_p(Fg.ptr.value + n * nk * 4), (ns + 2) * nk, scr.ptr)
    Fr = ref.array("F").reshape(6, ns + 2, nk)[:, n]
    Fs = Fg.get(np.float32).reshape(6, ns + 2, nk)[:, n]
    np.testing.assert_allclose(Fs, Fr, rtol=0, atol=2.5e-7 * np.abs(Fr).max())


def test_bunch_push_currents_and_window_shift_against_reference():
    """the RK4 push of a bunch in the fields of init, its currents, the window centroid and a shift by (3, -2)"""
    mk = lambda: bunch_deck(nx=32, ny=32, nz=64, nmacro=4096, laser=False)
    deck = mk()
    refpy.set_threads(1)
    ref = refpy.RefSim(mk())
    ref.init()
    tabs = Tables(g3d_amd.Sim(mk()))
    info = ref.info()
    N, ns, nk, np_ = info["N"], info["ns"], info["nk"], info["np"]
    s0, s1 = 0.0, deck.c.dz
    parts = g3d_amd.DeviceParts(ref.array("parts"))
    F = D(ref.array("F"))
    pc = tabs.push(deck, F, s1)
    ref.call("push", s0, s1)
    g3d_amd.kernel_call("g3d_bunch_push", C.byref(pc), C.byref(parts.c), 0, np_, s0)
    got, want = parts.records(), ref.array("parts")
    np.testing.assert_array_equal(got["alive"], want["alive"])
    particle_parity(got, want, deck, tol=1e-13)
    # currents of the pushed bunch: the deposit in the reference's order, then the spectra
    plan = g3d_amd.FftPlan(deck.c.nx, deck.c.ny)
    spec = tabs.spec(deck, plan, keep0=0)
    zax = tabs.zax(deck)
    jw = D(nbytes=3 * (ns + 3) * nk * 8)
    jk = [D(nbytes=(ns + 1) * nk * 16) for _ in range(3)]
    scr = D(nbytes=(ns + 1) * nk * 16)
    ref.call("currents")
    g3d_amd.kernel_call("g3d_bunch_currents", C.byref(spec), C.byref(zax), tabs.d["iA"].ptr, tabs.d["iD"].ptr,
                        info["dx"], info["dy"], deck.c.zeta_filter, 1, C.byref(parts.c), np_, jw.ptr, jk[0].ptr,
                        jk[1].ptr, jk[2].ptr, scr.ptr)
    np.testing.assert_array_equal(jw.get(np.float64), ref.array("jw"))
    for k, name in enumerate(("jkx", "jky", "jkz")):
        assert rel(jk[k].get(np.complex128), ref.array(name)) <= 1e-14, name
    # the window centroid: weighted sums over the live bunch particles
    out = np.zeros(3)
    g3d_amd.kernel_call("g3d_bunch_centroid", C.byref(parts.c), np_, info["X0"], info["Y0"],
                        out.ctypes.data_as(C.c_void_p))
    live = want[(want["alive"] == 1) & (want["sp"] <= 0)]
    exp = [(live["w"] * (live["x"] + info["X0"])).sum(), (live["w"] * (live["y"] + info["Y0"])).sum(), live["w"].sum()]
    np.testing.assert_allclose(out, exp, rtol=1e-12)
    # a shift of the window by (3, -2) cells: spectra, F and particles
    names = ("Rx", "Ry", "Ez", "Bz", "j0x", "j0y")
    arrs = [D(ref.array(c)) for c in names]
    counts = (C.c_long * 6)(*[ns, ns, ns, ns, ns + 1, ns + 1])
    ptrs = (C.c_void_p * 6)(*[a.ptr.value for a in arrs])
    Fs = D(nbytes=F.nbytes)
    ref.call("shift_window", 3, -2)
    g3d_amd.kernel_call("g3d_window_shift", C.byref(spec), 3, -2, info["dx"], info["dy"], ptrs, counts, 6, F.ptr,
                        Fs.ptr, C.byref(parts.c), np_)
    for a, c in zip(arrs, names):
        assert rel(a.get(np.complex128), ref.array(c)) <= 1e-14, c
    np.testing.assert_array_equal(F.get(np.float32), ref.array("F"))
    got, want = parts.records(), ref.array("parts")
    np.testing.assert_array_equal(got["x"], want["x"])
    np.testing.assert_array_equal(got["y"], want["y"])


def test_sponge_and_comoving_against_reference():
    """sponge_and_real_fields on random Hermitian spectra (sponge layers on, the bound field of the bunch taken out and
    added back), and comoving_field of the bunch's half-cell charge, through the kernel level"""
    mk = lambda: bunch_deck(nx=32, ny=32, nz=64, nmacro=4096, laser=False, sponge=4.0)
    deck = mk()
    refpy.set_threads(1)
    ref = refpy.RefSim(mk())
    ref.init()
    tabs = Tables(g3d_amd.Sim(mk()))
    info = ref.info()
    ns, nk, nx, ny = info["ns"], info["nk"], deck.c.nx, deck.c.ny
    plan = g3d_amd.FftPlan(nx, ny)
    spec = tabs.spec(deck, plan)
    rng = np.random.default_rng(11)
    arrs = {}
    for n in ("Rx", "Ry", "Lx", "Ly", "Ez", "Bz", "jkz"):
        planes = ns + 1 if n == "jkz" else ns
        a = np.fft.fft2(rng.standard_normal((planes, ny, nx))).reshape(-1)
        ref.set(n, a)
        arrs[n] = D(a)
    masks = {k: D(tabs.sim.array(k)) for k in ("maskx", "masky")}
    F = D(nbytes=6 * (ns + 2) * nk * 4)
    scr = D(nbytes=5 * ns * nk * 16)
    ref.call("sponge_and_real_fields")
    g3d_amd.kernel_call("g3d_sponge_real_fields", C.byref(spec), 1, masks["maskx"].ptr, masks["masky"].ptr,
                        info["beta_b"], info["gm2_b"], *[arrs[n].ptr for n in ("Rx", "Ry", "Lx", "Ly", "Ez", "Bz", "jkz")],
                        F.ptr, scr.ptr)
    for n in ("Rx", "Ry"):
        assert rel(arrs[n].get(np.complex128), ref.array(n)) <= 1e-13, n
    Fr = ref.array("F").reshape(6, ns + 2, nk)[:, :ns]
    Fg = F.get(np.float32).reshape(6, ns + 2, nk)[:, :ns]
    assert rel(Fg, Fr) <= 1.2e-7
    # the co-moving start field of the bunch's half-cell charge
    ref2 = ref                       # its bunch as init loaded it; currents recomputes jkz from it
    ref2.call("set_starting", 1)
    ref2.call("currents")
    jkz = D(ref2.array("jkz"))
    p = ref2.array("parts")
    assert len(p) > 0
    s0 = s1 = 0.0
    for q in p:
        if q["alive"] == 1:
            s0 += q["w"] * q["pz"] / math.sqrt(1 + q["px"] * q["px"] + q["py"] * q["py"] + q["pz"] * q["pz"])
            s1 += q["w"]
    beta = s0 / s1
    ref2.call("comoving_field")
    outs = {n: D(nbytes=ns * nk * 16) for n in ("Rx", "Ry", "Ez", "Bz")}
    scr2 = D(nbytes=2 * ns * nk * 16)
    g3d_amd.kernel_call("g3d_comoving", C.byref(spec), jkz.ptr, beta, 1 - beta * beta,
                        *[outs[n].ptr for n in ("Rx", "Ry", "Ez", "Bz")], scr2.ptr)
    assert ref2.info()["beta_b"] == beta
    for n in ("Rx", "Ry", "Ez", "Bz"):
        assert rel(outs[n].get(np.complex128), ref2.array(n)) <= 1e-13, n
