"""ctypes front of the C restatement (oracle/_build/libg3doracle.so) and the grid setup it needs, restated in numpy
from the reference's Simulation constructor (src/sim_setup.cpp:17-33 x modes, 63-82 Fourier y, 148-164 zeta).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's CPU baseline may use it."""
import ctypes as C
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "gem-pic-3d_amd"))
import g3d_amd  # noqa: E402

LIB_PATH = os.path.join(HERE, "_build", "libg3doracle.so")
D = C.c_double
PD = C.POINTER(C.c_double)


class Axis(C.Structure):
    _fields_ = [("X", PD), ("n", C.c_int), ("uniform", C.c_int), ("x0", D), ("ih", D)]


class Fields(C.Structure):
    _fields_ = [("F", C.POINTER(C.c_float)), ("a", C.c_int), ("ns", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
                ("yf", C.c_int), ("self", C.c_int), ("nk", C.c_long), ("dx", D), ("dy", D), ("Lx", D), ("Ly", D),
                ("zax", Axis), ("yax", Axis)]


class Laser(C.Structure):
    _fields_ = [("on", C.c_int), ("pol", C.c_int)] + [(n, D) for n in ("a0", "kL", "tau", "cep", "eta_c", "kx", "kz")]


class PushCtx(C.Structure):
    _fields_ = [("f", Fields), ("laser", Laser)] + \
               [(n, D) for n in ("X0", "Y0", "zc", "dz_push", "push_phase", "zhand", "zmax", "ylo", "yhi", "s1",
                                 "vz_min", "hmin")] + [("freeze", C.c_int)]


class Ion(C.Structure):
    _fields_ = [("mass", D), ("mobile", C.c_int), ("ionizing", C.c_int), ("Zmax", C.c_int), ("off", C.c_int),
                ("Emin", D)]


class Adk(C.Structure):
    _fields_ = [(n, D) for n in ("logpre", "power", "expo", "Epeak", "Emin")]


class StreamCtx(C.Structure):
    _fields_ = [("f", Fields), ("laser", Laser), ("X0", D), ("Y0", D), ("zc", D), ("ylo", D), ("yhi", D),
                ("zs", PD), ("J", PD), ("ia", PD), ("id", PD), ("Delta", D), ("s1", D), ("sTrap", D),
                ("vz_beam", D), ("n", C.c_int), ("ions", C.POINTER(Ion)), ("adk", C.POINTER(Adk)), ("born", C.c_void_p),
                ("bcnt", C.POINTER(C.c_int)), ("bcap", C.c_int), ("key", C.c_uint64)]


class MarchCtx(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("Rx", "Ry", "Ez", "Bz", "Lx", "Ly", "j0x", "j0y", "jkx", "jky", "jkz")] + \
               [(n, PD) for n in ("kx", "ky", "iD", "iA", "hcg", "wlo", "whi")] + \
               [("D", D), ("a", C.c_int), ("N", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nk", C.c_long)]


_lib = None


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise ImportError("oracle/_build/libg3doracle.so is not built (make -C oracle oracle)")
        L = C.CDLL(LIB_PATH)
        V = C.c_void_p
        sig = {
            "or_fft_many": ([C.c_int, C.c_int, C.c_int, V, C.c_int, C.c_int], None),
            "or_tsc_index": ([D, V, V], None), "or_axis_locate": ([V, D, V, V], None),
            "or_gather": ([V, D, D, D, V, V], None), "or_laser_field": ([V, D, D, D, V, V], None),
            "or_boris": ([V, V, V, D, D], None), "or_bunch_push": ([V, V, C.c_long, D], None),
            "or_bunch_deposit": ([V, C.c_long, V, V, V, D, D, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, V], None),
            "or_stream_halfcell": ([V, V, C.c_long], None), "or_adk_rate": ([V, D], D),
            "or_adk_levels": ([V, C.c_int, C.c_int, C.c_int, D, D, D, V], None),
            "or_mix64": ([C.c_uint64], C.c_uint64), "or_uniform01": ([C.c_uint64], D),
            "or_key_step": ([C.c_uint64, C.c_long], C.c_uint64), "or_key_slice": ([C.c_uint64, C.c_int, C.c_int], C.c_uint64),
            "or_key_particle": ([C.c_uint64, C.c_uint64], C.c_uint64), "or_march": ([V, C.c_int, C.c_int], None),
            "or_stream_spectra": ([C.c_int, C.c_int, V, V, V, V, V, D, C.c_int, V], None),
            "or_real_fields": ([C.c_int, C.c_int, C.c_int, V, V, V, V, V, V, V, C.c_long], None),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def dptr(a):
    return a.ctypes.data_as(PD)


# ---------------------------------------------------------------- grid setup (src/sim_setup.cpp)
class Grid:
    """k vectors, filters, sponge masks and zeta weights of a deck, as the reference's constructor sets them."""

    def __init__(self, deck):
        c = deck.c
        self.nx, self.ny, self.N = c.nx, c.ny, c.nzeta
        self.nk = c.nx * c.ny
        self.Lx, self.Ly = c.Lx, c.Ly
        self.dx = c.Lx / c.nx
        self.dy = c.Ly / c.ny
        nx, ny = c.nx, c.ny

        def prof(dd, w):
            if w <= 0 or dd >= w:
                return 1.0
            r = 1 - dd / w
            return 1 - c.sponge_rate * r * r

        self.kxs = np.array([2 * math.pi * (i if i <= nx // 2 else i - nx) / c.Lx for i in range(nx)])
        self.kx = np.array([0.0 if 2 * i == nx else self.kxs[i] for i in range(nx)])
        self.filt = np.array([0.0 if 2 * i == nx else (math.pow(math.cos(0.5 * self.kxs[i] * self.dx), 2)
                                                       if c.k_filter else 1.0) for i in range(nx)])
        self.maskx = np.array([prof(min(i * self.dx, c.Lx - i * self.dx), c.sponge) for i in range(nx)])
        self.kys = np.array([2 * math.pi * (j if j <= ny // 2 else j - ny) / c.Ly for j in range(ny)])
        self.ky = np.array([0.0 if 2 * j == ny else self.kys[j] for j in range(ny)])
        self.filty = np.array([0.0 if 2 * j == ny else (math.pow(math.cos(0.5 * self.kys[j] * self.dy), 2)
                                                        if c.k_filter else 1.0) for j in range(ny)])
        self.masky = np.array([prof(min(j * self.dy, c.Ly - j * self.dy), c.sponge_y) for j in range(ny)])
        self.iD = np.full(ny, 1 / self.dy)
        self.iA = np.full(ny, 1 / self.dy)
        self.freehand = not c.y_fourier
        self.ynode = None
        if self.freehand:                    # src/sim_setup.cpp:86-108: walls at the end nodes
            Y = np.array(deck.ynode, dtype=np.float64)
            Dl = np.diff(Y)
            A = np.empty(ny)
            A[0] = 0.5 * Dl[0]
            A[1:] = 0.5 * (Dl[:-1] + Dl[1:])
            self.iD, self.iA = 1 / Dl, 1 / A
            self.ynode = Y
            self.y_uniform = bool(np.all(Dl == Dl[0]))
            self.y_ih = 1 / Dl[0]
        Z = np.array(deck.zeta, dtype=np.float64)
        self.zeta = Z
        N = self.N
        hh = [Z[m + 1] - Z[m] for m in range(N - 1)]
        self.uniform = all(h == hh[0] for h in hh)
        if self.uniform:
            hh = [Z[-1] / (N - 1)] * (N - 1)
        self.hcg = np.array([hh[min(max(m, 0), N - 2)] for m in range(-2, N + 1)])
        self.wlo = np.array([self.hcg[n + 2] / (self.hcg[n + 1] + self.hcg[n + 2]) for n in range(N)])
        self.whi = np.array([self.hcg[n + 1] / (self.hcg[n + 1] + self.hcg[n + 2]) for n in range(N)])
        self.ih = 1 / hh[0]
        self._keep = []

    def zaxis(self):
        return Axis(dptr(self.zeta), self.N, int(self.uniform), 0.0, self.ih)

    def yaxis(self):
        if self.freehand:                    # the node table (src/sim_setup.cpp:105-108)
            return Axis(dptr(self.ynode), self.ny + 1, int(self.y_uniform), float(self.ynode[0]), self.y_ih)
        return Axis(None, self.ny + 1, 1, 0.0, 1 / self.dy)

    def fields(self, F, ns=None):
        ns = self.N if ns is None else ns
        self._keep.append(F)
        return Fields(F.ctypes.data_as(C.POINTER(C.c_float)), 0, ns, self.nx, self.ny, int(not self.freehand), 1,
                      self.nk, self.dx, self.dy, self.Lx, self.Ly, self.zaxis(), self.yaxis())


def laser_of(deck):
    c = deck.c
    pol = c.laser_pol
    return Laser(int(c.laser_on), pol, c.a0, c.kL, c.tau, c.cep, c.start_fwhm * c.tau, math.sin(c.theta),
                 math.cos(c.theta))


def mix64(z):
    return lib().or_mix64(z)


def fft_many(nx, ny, two_d, data, inverse):
    a = np.ascontiguousarray(data, dtype=np.complex128).copy()
    lib().or_fft_many(nx, ny, int(two_d), ptr(a), a.shape[0] if a.ndim == 3 else a.size // (nx * ny), int(inverse))
    return a


def march(g, arrays, D, lo=0, hi=None):
    """or_march on host copies: arrays holds Rx Ry Ez Bz Lx Ly (N planes) and j0x j0y jkx jky jkz (N+1 planes)"""
    hi = g.N if hi is None else hi
    a = {k: np.ascontiguousarray(v, dtype=np.complex128).copy() for k, v in arrays.items()}
    c = MarchCtx(*[ptr(a[k]) for k in ("Rx", "Ry", "Ez", "Bz", "Lx", "Ly", "j0x", "j0y", "jkx", "jky", "jkz")],
                 dptr(g.kx), None if g.freehand else dptr(g.ky), dptr(g.iD), dptr(g.iA), dptr(g.hcg), dptr(g.wlo),
                 dptr(g.whi),
                 D, 0, g.N, g.nx, g.ny, g.nk)
    lib().or_march(C.byref(c), lo, hi)
    return a


def stream_spectra(g, J, a2, keep0):
    out = np.zeros(3 * g.nk, dtype=np.complex128)
    J = np.ascontiguousarray(J, dtype=np.float64)
    lib().or_stream_spectra(g.nx, g.ny, ptr(J), ptr(g.kxs), ptr(g.kys), ptr(g.filt), ptr(g.filty), a2, int(keep0),
                            ptr(out))
    return out


def ion_tables(deck):
    """the per-species constants and the ADK levels of the deck's ion blocks, as setup_ions forms them
    (src/sim_ions.cpp:68-87): code units of the deck's wavelength, T = twice the window length"""
    c = deck.c
    Lu = c.lambda0 / (2 * math.pi)
    omega0 = g3d_amd.C_SI / Lu
    E0 = g3d_amd.ME_SI * g3d_amd.C_SI * omega0 / g3d_amd.E_SI
    T = 2 * float(deck.zeta[-1])
    ions, levels = [], []
    for k in range(c.n_ions):
        s = c.ions[k]
        off, Emin = len(levels), 1e300
        if s.ionization:
            L = adk_levels([s.ip_eV[i] for i in range(s.n_ip)], s.Z0, s.Zmax, E0, omega0, T)
            Emin = min(float(L[z][4]) for z in range(s.Z0, s.Zmax))
            levels.extend(L)
        ions.append((s.mass, int(s.mobile), int(s.ionization), s.Zmax, off, Emin))
    return ions, np.array(levels).reshape(-1, 5)


def stream_halfcell(g, deck, F, parts, n, s0, s1, X0, Y0, J=None, key=0, bcap=0):
    """one streaming pass through half-cell n in list order (Fourier or free-hand y; with the deck's ion species,
    their ADK levels and the draw key, room for bcap freed electrons); returns (particles, J, freed electrons,
    freed-electron count)"""
    c = deck.c
    p = np.ascontiguousarray(parts, dtype=g3d_amd.PARTICLE).copy()
    J = np.zeros(5 * g.nk) if J is None else J
    F = np.ascontiguousarray(F, dtype=np.float32)
    zs = np.ascontiguousarray(g.zeta)
    cnt = C.c_int(0)
    born = np.zeros(max(bcap, 1), dtype=g3d_amd.PARTICLE)
    sc = StreamCtx()
    sc.f = g.fields(F)
    sc.laser = laser_of(deck)
    ylo, yhi = (float(g.ynode[1]), float(g.ynode[-2])) if g.freehand else (-1e300, 1e300)
    sc.X0, sc.Y0, sc.zc, sc.ylo, sc.yhi = X0, Y0, c.zeta_center, ylo, yhi
    sc.zs = dptr(zs); sc.J = dptr(J); sc.ia = dptr(g.iA); sc.id = dptr(g.iD)
    sc.Delta = s1 - s0; sc.s1 = s1; sc.sTrap = s1 + (c.trap_advance - 1) * (s1 - s0); sc.vz_beam = c.vz_beam
    sc.n = n; sc.bcnt = C.pointer(cnt); sc.bcap = bcap; sc.key = key
    ions, levels = ion_tables(deck)
    ion_arr = (Ion * max(1, len(ions)))(*[Ion(*t) for t in ions])
    adk_arr = (Adk * max(1, len(levels)))(*[Adk(*map(float, L)) for L in levels])
    sc.ions = C.cast(ion_arr, C.POINTER(Ion))
    sc.adk = C.cast(adk_arr, C.POINTER(Adk))
    sc.born = born.ctypes.data_as(C.c_void_p)
    lib().or_stream_halfcell(C.byref(sc), ptr(p), len(p))
    return p, J, born[:min(cnt.value, bcap)].copy(), cnt.value


def bunch_deposit(g, parts):
    """the beam deposit jw [3][N+3][nk] of a particle list on one block (src/sim_particles.cpp:66-89)"""
    p = np.ascontiguousarray(parts, dtype=g3d_amd.PARTICLE)
    J = np.zeros(3 * (g.N + 3) * g.nk)
    za = g.zaxis()
    lib().or_bunch_deposit(ptr(p), len(p), C.byref(za), ptr(g.hcg), ptr(g.iA), g.dx, g.dy, g.nx, g.ny, 0, g.N, g.N,
                           ptr(J))
    return J


def key_chain(seed, step, n, pass_, ids):
    """the draw bases ion_key_particle(ion_key_slice(ion_key_step(seed, step), n, pass), id) (src/ionization.hpp:44-46)"""
    L = lib()
    ks = L.or_key_step(seed, step)
    kl = L.or_key_slice(ks, n, pass_)
    return ks, kl, [L.or_key_particle(kl, int(i)) for i in ids]


def bunch_push(g, deck, F, parts, s0, s1, X0, Y0, plasma=False):
    c = deck.c
    p = np.ascontiguousarray(parts, dtype=g3d_amd.PARTICLE).copy()
    F = np.ascontiguousarray(F, dtype=np.float32)
    pc = PushCtx()
    pc.f = g.fields(F)
    pc.laser = laser_of(deck)
    pc.X0, pc.Y0, pc.zc, pc.dz_push, pc.push_phase = X0, Y0, c.zeta_center, c.dz_push or c.dz, c.push_phase
    pc.zhand, pc.zmax, pc.ylo, pc.yhi, pc.s1 = -1e300, g.zeta[-1], -1e300, 1e300, s1
    pc.vz_min = c.vz_min if plasma else 0.0
    pc.hmin = float(np.min(np.diff(g.zeta)))
    pc.freeze = int(c.freeze_bunch)
    lib().or_bunch_push(C.byref(pc), ptr(p), len(p), s0)
    return p


def real_fields(g, arrays, count):
    fpl = (count + 2) * g.nk
    F = np.zeros(6 * fpl, dtype=np.float32)
    a = [np.ascontiguousarray(arrays[k], dtype=np.complex128) for k in ("Rx", "Ry", "Lx", "Ly", "Ez", "Bz")]
    lib().or_real_fields(g.nx, g.ny, count, *[ptr(x) for x in a], ptr(F), fpl)
    return F


def adk_levels(ip, Z0, Zmax, E0, omega0, T):
    out = (Adk * Zmax)()
    ipa = np.ascontiguousarray(ip, dtype=np.float64)
    lib().or_adk_levels(ptr(ipa), len(ipa), Z0, Zmax, E0, omega0, T, C.cast(out, C.c_void_p))
    return np.array([[L.logpre, L.power, L.expo, L.Epeak, L.Emin] for L in out])


def adk_rate(level, E):
    L = Adk(*level)
    return lib().or_adk_rate(C.byref(L), E)
