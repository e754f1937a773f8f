"""Python view of the MI355X march library (libg3d.so) over its C ABI (include/g3d.h).

The configuration helpers mirror the reference's Config grid helpers (src/sim_config.cpp:5-47:
base_units, zeta_uniform, y_uniform) and its test defaults (tests/test_util.hpp:36-43), so that a run is set up
here as the reference's tests set it up.  `Sim` mirrors the reference's Simulation: init(), step(), the member
functions by name (call), and host copies of the device arrays in the reference's layouts.

The library is loaded from this directory (lib/libg3d.so); if it is missing the import fails with an error -
there is no CPU fallback.
"""
import bisect
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libg3d.so")

# SI constants of the reference's units (src/sim_internal.hpp:10)
C_SI, E_SI, ME_SI, EPS0 = 299792458.0, 1.602176634e-19, 9.1093837015e-31, 8.8541878128e-12


class Glaser(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("a0", "kL", "waist", "tau", "zeta_c", "z_focus", "cep", "pol_angle")] + \
               [("circular", C.c_int), ("helicity", C.c_int), ("n_env", C.c_int), ("env_z0", C.c_double),
                ("env_dz", C.c_double), ("env", C.c_double * 32)]


class Bunch(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("charge", "gamma", "sigx", "sigy", "sigz", "emitx", "emity", "espread",
                                          "zeta_center")] + \
               [("nmacro", C.c_long), ("seed", C.c_ulong), ("name", C.c_char * 16)]


class Ion(C.Structure):
    _fields_ = [("mass", C.c_double), ("atoms", C.c_double), ("Z0", C.c_int), ("Zmax", C.c_int), ("mobile", C.c_int),
                ("ionization", C.c_int), ("n_ip", C.c_int), ("ip_eV", C.c_double * 16), ("name", C.c_char * 16)]


class Config(C.Structure):
    _fields_ = [
        ("lambda0", C.c_double), ("nx", C.c_int), ("ny", C.c_int), ("nzeta", C.c_int), ("Lx", C.c_double),
        ("Ly", C.c_double), ("y_fourier", C.c_int), ("zeta", C.POINTER(C.c_double)), ("ynode", C.POINTER(C.c_double)),
        ("dz", C.c_double), ("z_end", C.c_double), ("dz_push", C.c_double), ("sponge", C.c_double),
        ("sponge_y", C.c_double), ("sponge_rate", C.c_double), ("move_x", C.c_int), ("move_y", C.c_int),
        ("window_lag", C.c_int), ("self_fields", C.c_int), ("zeta_filter", C.c_int), ("k_filter", C.c_int),
        ("push_phase", C.c_double), ("laser_on", C.c_int), ("laser_pol", C.c_int), ("a0", C.c_double),
        ("kL", C.c_double), ("theta", C.c_double), ("tau", C.c_double), ("cep", C.c_double),
        ("start_fwhm", C.c_double), ("stop_fwhm", C.c_double), ("charge", C.c_double), ("gamma", C.c_double),
        ("sigx", C.c_double), ("sigy", C.c_double), ("sigz", C.c_double), ("emitx", C.c_double),
        ("emity", C.c_double), ("espread", C.c_double), ("zeta_center", C.c_double), ("nmacro", C.c_long),
        ("seed", C.c_ulong), ("n_prof", C.c_int), ("prof_zmin", C.c_double), ("prof_zmax", C.c_double),
        ("prof", C.c_double * 16), ("n_bunch_more", C.c_int), ("bunch_more", Bunch * 3), ("n_glasers", C.c_int), ("glasers", Glaser * 4), ("plasma_on", C.c_int),
        ("n0", C.c_double), ("r_max", C.c_double), ("ppc_x", C.c_int), ("ppc_y", C.c_int),
        ("smooth_length", C.c_double), ("n_pz", C.c_int), ("pz_points", C.c_double * 16),
        ("pz_values", C.c_double * 16), ("parabolic", C.c_int), ("channel_depth", C.c_double),
        ("r_channel", C.c_double), ("iterations", C.c_int), ("vz_min", C.c_double), ("vz_beam", C.c_double),
        ("trap_advance", C.c_double), ("predictor_extrapolate", C.c_int), ("free_electrons", C.c_int),
        ("n_ions", C.c_int), ("ions", Ion * 4), ("freeze_bunch", C.c_int), ("start_relax", C.c_int),
        ("box_mean", C.c_int), ("mixed", C.c_int)]


class SimInfo(C.Structure):
    _fields_ = [(n, C.c_long) for n in ("N", "ns", "nk", "np", "nstrm", "istep", "born_cap")] + \
               [(n, C.c_double) for n in ("s", "dx", "dy", "X0", "Y0", "beta_b", "gm2_b")] + \
               [(n, C.c_long) for n in ("trapped_total", "restreamed_total", "lost_total", "pushes_stream",
                                         "pushes_beam")]


# the 104-byte particle record (src/sim.hpp:217)
PARTICLE = np.dtype([("x", "f8"), ("y", "f8"), ("zeta", "f8"), ("px", "f8"), ("py", "f8"), ("pz", "f8"), ("w", "f8"),
                     ("rem", "f8"), ("z", "f8"), ("q", "f8"), ("qm", "f8"), ("alive", "i4"), ("sp", "i4"),
                     ("id", "u8")])
assert PARTICLE.itemsize == 104


class Deck:
    """A run configuration in code units with the reference's defaults (src/sim.hpp:36-187).

    The grid tables live here (numpy arrays) so the C struct can borrow them."""

    def __init__(self, lambda0=10e-6):
        self.c = Config()
        c = self.c
        c.lambda0 = lambda0
        c.dz_push = 0.0
        c.sponge_rate = 0.1
        c.move_x = 1; c.move_y = 1; c.window_lag = 4
        c.self_fields = 1; c.zeta_filter = 1; c.k_filter = 1
        c.push_phase = 0.05
        c.kL = 1.0; c.tau = 1.0; c.start_fwhm = 3.5; c.stop_fwhm = 3.5
        c.gamma = 1.0; c.seed = 1
        c.r_max = 1e300; c.r_channel = 1.0
        c.ppc_x = 2; c.ppc_y = 2; c.iterations = 2
        c.vz_min = 0.8; c.vz_beam = 0.9; c.trap_advance = 1.0
        c.predictor_extrapolate = 1; c.free_electrons = 1
        c.start_relax = 1; c.box_mean = 1
        self.zeta = None
        self.ynode = None

    # units (src/sim_config.cpp:5-11)
    @property
    def Lu(self):
        return self.c.lambda0 / (2 * math.pi)

    @property
    def nc(self):
        w0 = C_SI / self.Lu
        return EPS0 * ME_SI * w0 * w0 / (E_SI * E_SI)

    def zeta_uniform(self, n, Lw):                     # src/sim_config.cpp:14-20
        h = Lw / (n - 1)
        self.zeta = np.array([i * h for i in range(n)], dtype=np.float64)
        self.c.nzeta = n
        self.c.zeta = self.zeta.ctypes.data_as(C.POINTER(C.c_double))

    def y_uniform(self, n, Ly):                        # src/sim_config.cpp:32-38
        h = Ly / n
        self.ynode = np.array([(j - n // 2) * h for j in range(n + 1)], dtype=np.float64)
        self.c.ny = n; self.c.Ly = Ly; self.c.y_fourier = 1
        self.c.ynode = self.ynode.ctypes.data_as(C.POINTER(C.c_double))

    def zeta_table(self, pts, steps, Lw):               # src/sim_config.cpp:21-30
        m = StepMapping(pts, steps)
        cells = int(math.ceil(m.index(Lw) - 1e-9))
        if cells < 3:
            raise ValueError("grid: the zeta window needs at least 3 cells")
        self.zeta = np.array([m.coord(i) for i in range(cells + 1)], dtype=np.float64)
        self.c.nzeta = cells + 1
        self.c.zeta = self.zeta.ctypes.data_as(C.POINTER(C.c_double))

    def y_table(self, pts, steps, ymax):               # src/sim_config.cpp:39-48
        """free-hand y: nodes symmetric about y = 0 from the step table, conducting walls at the ends"""
        m = StepMapping(pts, steps)
        half = int(math.ceil(m.index(ymax) - 1e-9))
        if half < 2:
            raise ValueError("grid: the y grid needs at least 2 cells on each side")
        ny = 2 * half
        y = np.zeros(ny + 1)
        for k in range(half + 1):
            y[half + k] = m.coord(k)
            y[half - k] = -m.coord(k)
        self.y_nodes(y)

    def y_nodes(self, y):
        """free-hand y from an explicit node table (strictly increasing; the ends are the walls)"""
        y = np.array(y, dtype=np.float64)
        if y.ndim != 1 or len(y) < 5 or not np.all(np.diff(y) > 0):
            raise ValueError("free-hand y: at least 5 strictly increasing nodes")
        self.ynode = y
        self.c.ny = len(y) - 1; self.c.Ly = float(y[-1] - y[0]); self.c.y_fourier = 0; self.c.move_y = 0
        self.c.ynode = self.ynode.ctypes.data_as(C.POINTER(C.c_double))

    def grid(self, nx, ny, nz, Lx, Ly, Lw):
        """Fourier x and y, nz uniform slices; no sponge, fixed window (tests/test_util.hpp:36-43)."""
        self.c.nx = nx; self.c.Lx = Lx
        self.y_uniform(ny, Ly)
        self.zeta_uniform(nz, Lw)
        self.c.sponge = self.c.sponge_y = 0.0
        self.c.move_x = 0
        return self

    def add_grid_laser(self, a0, waist, tau, zeta_c, z_focus=0.0, kL=1.0, cep=0.0, pol_angle=0.0):
        g = self.c.glasers[self.c.n_glasers]
        g.a0 = a0; g.kL = kL; g.waist = waist; g.tau = tau; g.zeta_c = zeta_c; g.z_focus = z_focus; g.cep = cep
        g.pol_angle = pol_angle; g.circular = 0; g.helicity = 1
        self.c.n_glasers += 1
        return self

    def set_envelope(self, k, values, z0, dz):
        """grid laser k: the piecewise-linear envelope a(zeta) through values at z0 + i dz in place of the Gaussian
        of tau (the reference's envelope_function as a sum of hat functions)"""
        if not 1 <= len(values) <= 32:
            raise ValueError("an envelope has 1 ... 32 points")
        g = self.c.glasers[k]
        g.n_env = len(values); g.env_z0 = z0; g.env_dz = dz
        for i, v in enumerate(values):
            g.env[i] = v
        return self

    def add_bunch(self, name, charge, gamma, sigx, sigy, sigz, zeta_center, nmacro, seed, emitx=0.0, emity=0.0,
                  espread=0.0):
        """a further Gaussian electron bunch (the reference's bunch2, bunch3, ...; species -1, -2, -3), in code
        units; the first bunch is the deck's single-bunch fields"""
        if self.c.n_bunch_more >= 3:
            raise ValueError("at most 4 bunches")
        b = self.c.bunch_more[self.c.n_bunch_more]
        b.charge, b.gamma, b.sigx, b.sigy, b.sigz = charge, gamma, sigx, sigy, sigz
        b.emitx, b.emity, b.espread, b.zeta_center = emitx, emity, espread, zeta_center
        b.nmacro, b.seed, b.name = nmacro, seed, name.encode()[:15]
        self.c.n_bunch_more += 1
        return self

    def set_profile(self, values, zmin, zmax):
        """the bunch's current profile along zeta: the piecewise-linear profile_z through values at evenly spaced
        points from zmin to zmax (the reference's distribution_z = function), sampled by its inverse CDF"""
        if not 2 <= len(values) <= 16:
            raise ValueError("a profile has 2 ... 16 points")
        c = self.c
        c.n_prof = len(values); c.prof_zmin = zmin; c.prof_zmax = zmax
        for i, v in enumerate(values):
            c.prof[i] = v
        c.zeta_center = 0.5 * (zmin + zmax)
        return self

    def add_ion(self, mass, Z0, Zmax, atoms=1.0, mobile=True, ionization=False, ip_eV=(), name=""):
        """an ion species (src/sim.hpp IonSpecies): mass in electron masses, atoms per plasma electron, charge
        states Z0 .. Zmax, ionization potentials [eV] of the levels Z0 ... Zmax-1 (indexed by charge state)"""
        if self.c.n_ions >= 4:
            raise ValueError("at most 4 ion species")
        s = self.c.ions[self.c.n_ions]
        s.mass = mass; s.atoms = atoms; s.Z0 = Z0; s.Zmax = Zmax; s.mobile = int(mobile); s.ionization = int(ionization)
        s.name = name.encode()[:15]
        s.n_ip = len(ip_eV)
        for k, v in enumerate(ip_eV):
            s.ip_eV[k] = v
        self.c.n_ions += 1
        return self

    def set(self, **kw):
        for k, v in kw.items():
            if not hasattr(self.c, k):
                raise AttributeError("config has no field " + k)
            setattr(self.c, k, v)
        if "dz" in kw and "dz_push" not in kw and self.c.dz_push == 0:
            self.c.dz_push = kw["dz"]
        return self

    def ptr(self):
        if self.c.dz_push == 0:
            self.c.dz_push = self.c.dz
        return C.byref(self.c)


class StepMapping:
    """A grid from a table of steps h_k at reference positions x_k >= 0 (src/grid.cpp:11-52): h(x) is linear
    between the points and constant beyond them, the index is xi(x) = int_0^x dx'/h(x') and the nodes are
    X(i) = xi^-1(i).  On a segment with slope b, xi = xi_k + ln(1 + b d/h_k)/b and X = x_k + h_k (e^{b t} - 1)/b."""

    def __init__(self, x, h):
        x, h = [float(v) for v in x], [float(v) for v in h]
        if not x or len(x) != len(h):
            raise ValueError("needs as many steps as reference points (at least one)")
        for k in range(len(x)):
            if not x[k] >= 0 or (k and not x[k] > x[k - 1]) or not h[k] > 0:
                raise ValueError("reference points >= 0 and increasing, steps > 0")
        if x[0] > 0:
            x, h = [0.0] + x, [h[0]] + h
        K = len(x)
        self.x, self.h, self.b, self.xi = x, h, [0.0] * K, [0.0] * K
        for k in range(K - 1):
            L = x[k + 1] - x[k]
            self.b[k] = (h[k + 1] - h[k]) / L
            self.xi[k + 1] = self.xi[k] + (L / h[k]) * _log1p_over(self.b[k] * L / h[k])

    def index(self, x):
        if x <= 0:
            return x / self.h[0]
        k = min(bisect.bisect_right(self.x, x) - 1, len(self.x) - 1)
        d = x - self.x[k]
        return self.xi[k] + (d / self.h[k]) * _log1p_over(self.b[k] * d / self.h[k])

    def coord(self, xi):
        if xi <= 0:
            return xi * self.h[0]
        k = min(bisect.bisect_right(self.xi, xi) - 1, len(self.xi) - 1)
        t = xi - self.xi[k]
        return self.x[k] + self.h[k] * t * _expm1_over(self.b[k] * t)


def _log1p_over(y):      # ln(1 + y)/y, its limit 1 - y/2 near 0
    return 1.0 - 0.5 * y if abs(y) < 1e-8 else math.log1p(y) / y


def _expm1_over(y):      # (e^y - 1)/y
    return 1.0 + 0.5 * y if abs(y) < 1e-8 else math.expm1(y) / y


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("libg3d.so is not built (%s): run __graft_entry__.build()" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sig = {
        "g3d_init": ([C.c_int], C.c_int), "g3d_last_error": ([], C.c_char_p), "g3d_sync": ([], C.c_int),
        "g3d_status_string": ([C.c_int], C.c_char_p),
        "g3d_malloc": ([C.POINTER(P), C.c_size_t], C.c_int), "g3d_free": ([P], C.c_int),
        "g3d_memset0": ([P, C.c_size_t], C.c_int),
        "g3d_memcpy_h2d": ([P, P, C.c_size_t], C.c_int), "g3d_memcpy_d2h": ([P, P, C.c_size_t], C.c_int),
        "g3d_memcpy_d2d": ([P, P, C.c_size_t], C.c_int),
        "g3d_fft_plan": ([C.c_int, C.c_int, C.c_int, C.POINTER(P)], C.c_int),
        "g3d_fft_exec": ([P, P, C.c_int, C.c_int], C.c_int), "g3d_fft_destroy": ([P], C.c_int),
        "g3d_march": ([P, C.c_int, C.c_int], C.c_int), "g3d_march_slice": ([P, C.c_int, C.c_int], C.c_int),
        "g3d_march_adjoint": ([P, C.c_int, C.c_int], C.c_int),
        "g3d_tsc_deposit": ([P, P, P], C.c_int), "g3d_tsc_gather": ([P, P, P, P, P], C.c_int),
        "g3d_boris": ([P], C.c_int), "g3d_boris_vjp": ([P, P], C.c_int),
        "g3d_stream_push": ([P], C.c_int), "g3d_stream_push_vjp": ([P, P], C.c_int),
        "g3d_bunch_push_piece": ([P], C.c_int), "g3d_bunch_push_vjp": ([P, P], C.c_int),
        "g3d_bunch_push": ([P, P, C.c_long, C.c_long, C.c_double], C.c_int),
        "g3d_stream_halfcell": ([P, P, P, C.c_long], C.c_int),
        "g3d_stream_spectra": ([P, P, P, P], C.c_int), "g3d_stream_spectra_static": ([P, P, P], C.c_int),
        "g3d_real_fields_slice": ([P, P, P, P, P, P, P, P, C.c_long, P], C.c_int),
        "g3d_bunch_currents": ([P, P, P, P, C.c_double, C.c_double, C.c_int, C.c_int, P, C.c_long, P, P, P, P, P],
                               C.c_int),
        "g3d_bunch_centroid": ([P, C.c_long, C.c_double, C.c_double, P], C.c_int),
        "g3d_sponge_real_fields": ([P, C.c_int, P, P, C.c_double, C.c_double, P, P, P, P, P, P, P, P, P], C.c_int),
        "g3d_comoving": ([P, P, C.c_double, C.c_double, P, P, P, P, P], C.c_int),
        "g3d_window_shift": ([P, C.c_int, C.c_int, C.c_double, C.c_double, P, P, C.c_int, P, P, P, C.c_long],
                             C.c_int),
        "g3d_sim_create": ([P, C.POINTER(P)], C.c_int), "g3d_sim_destroy": ([P], C.c_int),
        "g3d_sim_info_get": ([P, P], C.c_int), "g3d_sim_init": ([P], C.c_int), "g3d_sim_step": ([P], C.c_int),
        "g3d_sim_call": ([P, C.c_char_p, C.POINTER(C.c_double), C.c_int], C.c_int),
        "g3d_sim_array_size": ([P, C.c_char_p, C.POINTER(C.c_size_t)], C.c_int),
        "g3d_sim_get": ([P, C.c_char_p, P, C.c_size_t], C.c_int),
        "g3d_sim_set": ([P, C.c_char_p, P, C.c_size_t], C.c_int),
        "g3d_sim_set_particles": ([P, P, C.c_long], C.c_int), "g3d_sim_set_joining": ([P, P, C.c_long], C.c_int),
        "g3d_sim_get_joining": ([P, P, C.c_long, C.POINTER(C.c_long)], C.c_int),
        "g3d_sim_timing": ([P, C.c_int], C.c_int), "g3d_sim_timing_every": ([P, C.c_int], C.c_int),
        "g3d_sim_write_bunches": ([P, C.c_char_p, C.c_int], C.c_int),
        "g3d_sim_write_slices": ([P, C.c_char_p, C.c_double], C.c_int),
        "g3d_sim_diagnostics": ([P, C.c_char_p, C.c_int], C.c_int),
        "g3d_sim_field_energy": ([P, C.POINTER(C.c_double)], C.c_int),
        "g3d_sim_dump_particles": ([P, C.c_char_p], C.c_int),
        "g3d_sim_write_frame": ([P, C.c_char_p, C.c_int], C.c_int),
        "g3d_sim_write_axis": ([P, C.c_char_p], C.c_int),
        "g3d_sim_write_checkpoint": ([P, C.c_char_p], C.c_int),
        "g3d_sim_read_checkpoint": ([P, C.c_char_p, C.c_long], C.c_int),
        "g3d_sim_gradient": ([P, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int,
                              C.POINTER(C.c_double), C.c_int], C.c_int),
        "g3d_sim_tangent": ([P, C.c_int, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double), C.c_int,
                             C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
        "g3d_sim_timers": ([P, C.POINTER(C.c_double), C.POINTER(C.c_long), C.c_int, C.c_char_p, C.c_int], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
        if _lib.g3d_init(0) != 0:
            raise RuntimeError("g3d_init: " + _lib.g3d_last_error().decode())
    return _lib


def check(rc):
    if rc != 0:
        raise RuntimeError(lib().g3d_last_error().decode() or lib().g3d_status_string(rc).decode())


ARRAY_DTYPES = {"F": np.float32, "jw": np.float64, "jsr": np.float64, "rho_lines": np.float64, "rho_lines_i": np.float64, "kx": np.float64,
                "ky": np.float64, "hcg": np.float64, "wlo": np.float64, "whi": np.float64, "kxs": np.float64,
                "kys": np.float64, "filt": np.float64, "filty": np.float64, "zs": np.float64, "iA": np.float64,
                "iD": np.float64, "wsl": np.float64, "maskx": np.float64, "masky": np.float64, "ion_stats0": np.float64, "ion_stats1": np.float64,
                "ion_stats2": np.float64, "ion_stats3": np.float64, "laser": np.float64,
                "parts": PARTICLE, "strm": PARTICLE, "born": PARTICLE}


class SimBase:
    """Common front of the engine and of the reference harness (same method names)."""

    def array(self, name):
        n = self._size(name)
        dt = ARRAY_DTYPES.get(name, np.complex128)
        out = np.empty(n // np.dtype(dt).itemsize, dtype=dt)
        self._get(name, out)
        return out

    def info(self):
        i = SimInfo()
        self._info(i)
        return {f: getattr(i, f) for f, _ in SimInfo._fields_}


class Sim(SimBase):
    """The MI355X engine (g3d_sim)."""

    def __init__(self, deck):
        L = lib()
        self._deck = deck
        h = C.c_void_p()
        check(L.g3d_sim_create(deck.ptr(), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            check(lib().g3d_sim_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self):
        check(lib().g3d_sim_init(self.h))

    def step(self):
        check(lib().g3d_sim_step(self.h))

    def call(self, op, *args):
        a = (C.c_double * max(1, len(args)))(*args)
        check(lib().g3d_sim_call(self.h, op.encode(), a, len(args)))

    def _size(self, name):
        b = C.c_size_t()
        check(lib().g3d_sim_array_size(self.h, name.encode(), C.byref(b)))
        return b.value

    def _get(self, name, out):
        check(lib().g3d_sim_get(self.h, name.encode(), out.ctypes.data_as(C.c_void_p), out.nbytes))

    def set(self, name, arr):
        arr = np.ascontiguousarray(arr)
        check(lib().g3d_sim_set(self.h, name.encode(), arr.ctypes.data_as(C.c_void_p), arr.nbytes))

    def set_particles(self, p):
        p = np.ascontiguousarray(p, dtype=PARTICLE)
        check(lib().g3d_sim_set_particles(self.h, p.ctypes.data_as(C.c_void_p), len(p)))

    def set_joining(self, p):
        p = np.ascontiguousarray(p, dtype=PARTICLE)
        check(lib().g3d_sim_set_joining(self.h, p.ctypes.data_as(C.c_void_p), len(p)))

    def joining(self, cap=1 << 22):
        out = np.empty(cap, dtype=PARTICLE)
        n = C.c_long()
        check(lib().g3d_sim_get_joining(self.h, out.ctypes.data_as(C.c_void_p), cap, C.byref(n)))
        return out[:n.value].copy()

    def _info(self, i):
        check(lib().g3d_sim_info_get(self.h, C.byref(i)))

    # output and checkpoints in the reference's formats (src/sim_bunch_diag.cpp, src/sim_output.cpp,
    # src/sim_checkpoint.cpp)
    def write_bunches(self, path, header=False):
        check(lib().g3d_sim_write_bunches(self.h, str(path).encode(), int(header)))

    def write_slices(self, path, slice_length):
        check(lib().g3d_sim_write_slices(self.h, str(path).encode(), float(slice_length)))

    def diagnostics(self, path, header=False):
        check(lib().g3d_sim_diagnostics(self.h, str(path).encode(), int(header)))

    def field_energy(self):
        e = C.c_double()
        check(lib().g3d_sim_field_energy(self.h, C.byref(e)))
        return e.value

    def dump_particles(self, path):
        check(lib().g3d_sim_dump_particles(self.h, str(path).encode()))

    def write_frame(self, path, comps=255):
        check(lib().g3d_sim_write_frame(self.h, str(path).encode(), int(comps)))

    def write_axis(self, path):
        check(lib().g3d_sim_write_axis(self.h, str(path).encode()))

    def write_checkpoint(self, directory):
        check(lib().g3d_sim_write_checkpoint(self.h, str(directory).encode()))

    def read_checkpoint(self, directory, step):
        check(lib().g3d_sim_read_checkpoint(self.h, str(directory).encode(), int(step)))

    def gradient(self, steps, every=5):
        """On a new simulation: init() and `steps` steps, then the reverse (g3d_sim_gradient).  Returns the mean p_z of
        the live bunch at the end and its derivatives in the envelope values of the first grid laser and in the
        bunch's current-profile values (numpy float64; empty where the deck has none)."""
        c = self._deck.c
        nenv = int(c.glasers[0].n_env) if c.n_glasers > 0 else 0
        nprof = int(c.n_prof)
        loss = C.c_double()
        genv, gprof = np.zeros(max(nenv, 1)), np.zeros(max(nprof, 1))
        dp = C.POINTER(C.c_double)
        check(lib().g3d_sim_gradient(self.h, int(steps), int(every), C.byref(loss), genv.ctypes.data_as(dp), nenv,
                                     gprof.ctypes.data_as(dp), nprof))
        return loss.value, genv[:nenv], gprof[:nprof]

    def tangent(self, steps, denv, dprof):
        """On a new simulation: the loss of gradient() and its derivative along (denv, dprof) by the tangent-linear
        run (g3d_sim_tangent).  Returns (loss, dloss)."""
        denv = np.ascontiguousarray(denv, dtype=np.float64)
        dprof = np.ascontiguousarray(dprof, dtype=np.float64)
        loss, dloss = C.c_double(), C.c_double()
        dp = C.POINTER(C.c_double)
        e = denv if denv.size else np.zeros(1)
        p = dprof if dprof.size else np.zeros(1)
        check(lib().g3d_sim_tangent(self.h, int(steps), e.ctypes.data_as(dp), int(denv.size), p.ctypes.data_as(dp),
                                    int(dprof.size), C.byref(loss), C.byref(dloss)))
        return loss.value, dloss.value

    FAMILIES = ("stream_halfcell", "pass_rows_fwd", "pass_cols", "pass_rows_inv", "unused", "beam_push", "currents",
                "sponge_real_fields")

    def timing(self, on):
        """HIP event timing: True = every kernel family, False = off, or a list of family names"""
        if on is True:
            mask = -1
        elif not on:
            mask = 0
        else:
            mask = sum(1 << self.FAMILIES.index(n) for n in on)
        check(lib().g3d_sim_timing(self.h, mask))

    def timing_every(self, every):
        """time one launch in `every` of the streaming kernel (its launches carry their own events)"""
        check(lib().g3d_sim_timing_every(self.h, int(every)))

    def timers(self):
        nf = 8
        ms = (C.c_double * nf)()
        cnt = (C.c_long * nf)()
        names = C.create_string_buffer(nf * 32)
        check(lib().g3d_sim_timers(self.h, ms, cnt, nf, names, 32))
        out = {}
        for k in range(nf):
            nm = names.raw[k * 32:(k + 1) * 32].split(b"\0")[0].decode()
            out[nm] = (ms[k], cnt[k])
        return out


# ---------------------------------------------------------------- kernel level (include/g3d.h): the reference's
# device contexts over caller-owned device buffers
class Axis(C.Structure):
    """src/grid.hpp:20-46 Axis: uniform nodes x0 + i/ih, or a device node table X[0 .. n-1]"""
    _fields_ = [("X", C.c_void_p), ("n", C.c_int), ("uniform", C.c_int), ("x0", C.c_double), ("ih", C.c_double)]


class PlaneWave(C.Structure):
    _fields_ = [("on", C.c_int), ("pol", C.c_int)] + \
               [(n, C.c_double) for n in ("a0", "kL", "tau", "cep", "eta_c", "kx", "kz")]


class PushCtx(C.Structure):
    """src/sim_kernels.hpp:10-17 PushCtx without the external elements"""
    _fields_ = [("F", C.c_void_p)] + [(n, C.c_int) for n in ("a", "ns", "N", "nx", "ny", "self", "yf")] + \
               [("nk", C.c_long)] + [(n, C.c_double) for n in ("dx", "dy", "Lx", "Ly")] + \
               [("zax", Axis), ("yax", Axis), ("laser", PlaneWave)] + \
               [(n, C.c_double) for n in ("X0", "Y0", "zc", "dz_push", "push_phase", "zhand", "zmax", "ylo", "yhi",
                                          "s1", "vz_min", "hmin")] + [("freeze", C.c_int)]


class Parts(C.Structure):
    """g3d_parts: a device particle list as structure of arrays"""
    _fields_ = [(n, C.c_void_p) for n in ("x", "y", "zeta", "px", "py", "pz", "w", "z", "q", "qm", "rem", "alive",
                                          "sp", "id")]


class StreamCtx(C.Structure):
    """src/sim_stream.hpp:20-29 StreamCtx"""
    _fields_ = [("pc", PushCtx), ("zs", C.c_void_p), ("J", C.c_void_p), ("ia", C.c_void_p), ("id", C.c_void_p)] + \
               [(n, C.c_double) for n in ("Delta", "s1", "sTrap", "vz_beam")] + [("n", C.c_int)] + \
               [("ions", C.c_void_p), ("adk", C.c_void_p), ("born", Parts), ("bcnt", C.c_void_p), ("bcap", C.c_int),
                ("key", C.c_uint64)]


class SpecCtx(C.Structure):
    """the spectral geometry of the constructor (src/sim_setup.cpp:5-60, 145-164), device tables"""
    _fields_ = [(n, C.c_int) for n in ("nx", "ny", "ns", "N", "a")] + [("nk", C.c_long)] + \
               [(n, C.c_void_p) for n in ("kx", "ky", "kxs", "kys", "filt", "filty", "hcg", "wlo", "whi")] + \
               [("a2", C.c_double), ("keep0", C.c_int), ("plan", C.c_void_p)]


class DevArray:
    """a device buffer from g3d_malloc holding a copy of a host array (or `nbytes` zero bytes)"""

    def __init__(self, host=None, nbytes=None):
        self.ptr = C.c_void_p()
        self.nbytes = int(host.nbytes if host is not None else nbytes)
        check(lib().g3d_malloc(C.byref(self.ptr), max(self.nbytes, 8)))
        if host is not None and self.nbytes:
            h = np.ascontiguousarray(host)
            check(lib().g3d_memcpy_h2d(self.ptr, h.ctypes.data_as(C.c_void_p), self.nbytes))

    def get(self, dtype, count=None):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize if count is None else count, dtype=dtype)
        if out.nbytes:
            check(lib().g3d_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes))
        return out

    def __del__(self):
        try:
            if self.ptr:
                lib().g3d_free(self.ptr)
                self.ptr = C.c_void_p()
        except Exception:
            pass


SOA_FIELDS = ("x", "y", "zeta", "px", "py", "pz", "w", "z", "q", "qm", "rem", "alive", "sp", "id")


class DeviceParts:
    """a particle list on the device as structure of arrays (g3d_parts), made from 104-byte records"""

    def __init__(self, records=None, n=None):
        recs = np.zeros(n, dtype=PARTICLE) if records is None else np.ascontiguousarray(records, dtype=PARTICLE)
        self.n = len(recs)
        self.arrays = {f: DevArray(np.ascontiguousarray(recs[f])) for f in SOA_FIELDS}
        self.c = Parts(**{f: self.arrays[f].ptr.value for f in SOA_FIELDS})

    def records(self):
        out = np.zeros(self.n, dtype=PARTICLE)
        for f in SOA_FIELDS:
            out[f] = self.arrays[f].get(PARTICLE[f], self.n)
        return out


class FftPlan:
    """FftMany(nx, ny, two_d) of the C ABI"""

    def __init__(self, nx, ny, two_d=True):
        self.h = C.c_void_p()
        check(lib().g3d_fft_plan(nx, ny, int(two_d), C.byref(self.h)))

    def __del__(self):
        try:
            if self.h:
                lib().g3d_fft_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass


def kernel_call(name, *args):
    """call a kernel-level entry point of the C ABI and raise on a nonzero status"""
    check(getattr(lib(), name)(*args))
