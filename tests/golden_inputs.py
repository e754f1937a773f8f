"""Seeded inputs of the golden vectors (shared by tests/golden/make_golden.py and the tests that replay them)."""
import numpy as np

import g3d_amd
from decks import LAMBDA0

MARCH_D = 2.5
STREAM_N = 9
# NIST ionization energies of nitrogen [eV] (the reference's atomic-data row N, src/atomic_data.cpp:36)
N_IP = np.array([14.53413, 29.60125, 47.4453, 77.4735, 97.8901, 552.06733, 667.04612])


def small_deck(smooth=0.0):
    d = g3d_amd.Deck(LAMBDA0).grid(8, 8, 12, 16.0, 16.0, 11 * 0.5)
    d.set(dz=5.0, z_end=100.0, plasma_on=1, n0=1e-2, ppc_x=1, ppc_y=1, move_y=0, smooth_length=smooth)
    return d


def small_fh_deck():
    """free-hand y: 14 cells between walls at +-6.6/k0, graded 0.4 -> 0.8/k0; 8 Fourier x modes, 12 slices"""
    d = g3d_amd.Deck(LAMBDA0)
    d.c.nx = 8; d.c.Lx = 16.0
    d.y_table([0.0, 2.0, 5.0], [0.4, 0.4, 0.8], 6.5)
    d.zeta_uniform(12, 11 * 0.5)
    d.set(dz=5.0, z_end=100.0, plasma_on=1, n0=1e-2, ppc_x=1, ppc_y=1, move_x=0, sponge=0.0, sponge_y=0.0)
    return d


def bunch_small_deck():
    d = g3d_amd.Deck(LAMBDA0).grid(8, 8, 16, 16.0, 16.0, 15 * 0.5)
    d.set(dz=20.0, z_end=100.0, laser_on=1, laser_pol=1, a0=1.0, theta=2e-3, tau=30.0, start_fwhm=0.0,
          stop_fwhm=0.0, zeta_center=3.75)
    return d


def fft_input():
    r = np.random.default_rng(11)
    return r.standard_normal((3, 16, 8)) + 1j * r.standard_normal((3, 16, 8))


def march_input(sim):
    r = np.random.default_rng(12)
    out = {}
    for n in ("Rx", "Ry", "Ez", "Bz", "Lx", "Ly", "j0x", "j0y", "jkx", "jky", "jkz"):
        m = sim._size(n) // 16
        out[n] = r.standard_normal(m) + 1j * r.standard_normal(m)
    return out


def jsr_input():
    return np.random.default_rng(13).uniform(-1, 1, 5 * 64)


def stream_input(sim):
    """random fields of moderate amplitude and 64 electrons in half-cell STREAM_N (some waiting, some off the
    slice top, some in a strong transverse kick)"""
    r = np.random.default_rng(14)
    F = (0.05 * r.standard_normal(sim._size("F") // 4)).astype(np.float32)
    P = np.zeros(64, dtype=g3d_amd.PARTICLE)
    z = np.linspace(0, 5.5, 12)
    P["x"] = r.uniform(0, 16, 64)
    P["y"] = r.uniform(0, 16, 64)
    P["zeta"] = z[STREAM_N + 1]
    P["zeta"][::7] = z[STREAM_N] + 0.3 * (z[STREAM_N + 1] - z[STREAM_N])
    P["px"] = 0.3 * r.standard_normal(64)
    P["py"] = 0.3 * r.standard_normal(64)
    P["pz"] = 0.2 * r.standard_normal(64)
    P["pz"][:3] = 30.0                       # fast: trapped in this pass
    P["w"] = 1e-4
    P["q"] = -1
    P["qm"] = -1
    P["sp"] = 2
    P["alive"] = 1
    P["alive"][5] = 6
    P["zeta"][5] = z[STREAM_N] + 0.5 * (z[STREAM_N + 1] - z[STREAM_N])
    P["alive"][6] = 6
    P["zeta"][6] = z[STREAM_N] - 0.1         # waiting below the half-cell: stays
    P["alive"][7] = 0
    P["z"] = 0.0
    P["id"] = np.arange(64, dtype=np.uint64) + (1 << 32)
    return F, P


def push_input(sim):
    r = np.random.default_rng(15)
    F = (0.01 * r.standard_normal(sim._size("F") // 4)).astype(np.float32)
    P = np.zeros(32, dtype=g3d_amd.PARTICLE)
    P["x"] = 8 + r.standard_normal(32)
    P["y"] = 8 + r.standard_normal(32)
    P["zeta"] = 3.75 + 0.5 * r.standard_normal(32)
    P["px"] = 1e-2 * r.standard_normal(32)
    P["py"] = 1e-2 * r.standard_normal(32)
    P["pz"] = 1000 * (1 + 1e-3 * r.standard_normal(32))
    P["pz"][0] = 3.0                          # slow: small sub-steps
    P["w"] = 1e-3
    P["q"] = -1
    P["qm"] = -1
    P["alive"] = 1
    return F, P


def units(lambda0=0.8e-6):
    Lu = lambda0 / (2 * np.pi)
    omega0 = g3d_amd.C_SI / Lu
    E0 = g3d_amd.ME_SI * g3d_amd.C_SI * omega0 / g3d_amd.E_SI
    return E0, omega0


def ref_adk_table():
    import ctypes as C
    import refpy
    E0, omega0 = units()
    out = np.zeros((7, 5))
    ip = np.ascontiguousarray(N_IP)
    refpy.check(refpy.lib().ref_adk_levels(ip.ctypes.data_as(C.c_void_p), 7, 0, 7, E0, omega0, 2 * 500.0,
                                           out.ctypes.data_as(C.c_void_p)))
    return out


AMU = 1822.888486


def ion_deck():
    """small_deck with two nitrogen species ionized by ADK from the neutral atom: species 0 mobile, species 1 fixed
    (immobile ionizing ions deposit their charge and ionize in place, src/sim_stream.hpp:115-122)"""
    d = small_deck()
    d.add_ion(14.0067 * AMU, 0, 7, atoms=0.01, mobile=True, ionization=True, ip_eV=N_IP)
    d.add_ion(14.0067 * AMU, 0, 7, atoms=0.01, mobile=False, ionization=True, ip_eV=N_IP)
    d.set(seed=20240611)
    return d


ION_KEY = (20240611, 3, STREAM_N, 1)       # seed, step, half-cell, pass of the draws of the ion pass


def ion_stream_input(sim):
    """fields of about 0.3 (a few 1e11 V/m at lambda0 = 2 pi um: the first nitrogen levels ionize within a
    half-cell), 24 electrons, 24 mobile and 16 fixed nitrogen ions of charge states 0 ... 4 in half-cell STREAM_N"""
    r = np.random.default_rng(16)
    F = (0.3 * r.standard_normal(sim._size("F") // 4)).astype(np.float32)
    ne, nm, nf = 24, 24, 16
    n = ne + nm + nf
    P = np.zeros(n, dtype=g3d_amd.PARTICLE)
    z = np.linspace(0, 5.5, 12)
    P["x"] = r.uniform(0, 16, n)
    P["y"] = r.uniform(0, 16, n)
    P["zeta"] = z[STREAM_N + 1]
    P["zeta"][::5] = z[STREAM_N] + 0.4 * (z[STREAM_N + 1] - z[STREAM_N])
    P["px"] = 0.1 * r.standard_normal(n)
    P["py"] = 0.1 * r.standard_normal(n)
    P["pz"] = 0.05 * r.standard_normal(n)
    P["w"] = 1e-4
    P["alive"] = 1
    P["z"] = 0.0
    P["id"] = (np.uint64(5) << np.uint64(32)) + np.arange(n, dtype=np.uint64) * np.uint64(7)
    P["q"][:ne] = -1
    P["qm"][:ne] = -1
    P["sp"][:ne] = 2
    q = r.integers(0, 5, nm + nf).astype(np.float64)
    P["q"][ne:] = q
    P["qm"][ne:] = q / (14.0067 * AMU)
    P["px"][ne:] *= 1e-3
    P["py"][ne:] *= 1e-3
    P["pz"][ne:] *= 1e-3
    P["sp"][ne:ne + nm] = 3
    P["sp"][ne + nm:] = 4
    cap = int(sum(7 - int(v) for v in q))
    return F, P, cap


def fh_stream_deck():
    """free-hand y (small_fh_deck's graded nodes between the walls at +-6.6/k0) and a free-hand zeta table of
    graded slices (0.4 -> 0.7/k0), with the oblique plane wave"""
    d = g3d_amd.Deck(LAMBDA0)
    d.c.nx = 8; d.c.Lx = 16.0
    d.y_table([0.0, 2.0, 5.0], [0.4, 0.4, 0.8], 6.5)
    d.zeta_table([0.0, 2.0, 5.0], [0.4, 0.4, 0.7], 6.0)
    d.set(dz=5.0, z_end=100.0, plasma_on=1, n0=1e-2, ppc_x=1, ppc_y=1, move_x=0, sponge=0.0, sponge_y=0.0,
          laser_on=1, laser_pol=1, a0=0.5, theta=2e-3, tau=30.0, start_fwhm=0.0, stop_fwhm=0.0, zeta_center=3.0)
    return d


FH_STREAM_N = 7


def fh_stream_input(sim, deck):
    """48 electrons in half-cell FH_STREAM_N of the zeta table, spread over the free-hand y nodes up to the walls
    (some leave through a wall in this pass), a few waiting below or inside the half-cell"""
    r = np.random.default_rng(17)
    F = (0.05 * r.standard_normal(sim._size("F") // 4)).astype(np.float32)
    z = deck.zeta
    yn = deck.ynode
    n = 48
    P = np.zeros(n, dtype=g3d_amd.PARTICLE)
    P["x"] = r.uniform(0, 16, n)
    P["y"] = r.uniform(yn[1], yn[-2], n)
    P["y"][:4] = [yn[1] + 0.01, yn[-2] - 0.01, yn[2], 0.5 * (yn[0] + yn[1]) + 0.4]
    P["zeta"] = z[FH_STREAM_N + 1]
    P["zeta"][::6] = z[FH_STREAM_N] + 0.3 * (z[FH_STREAM_N + 1] - z[FH_STREAM_N])
    P["px"] = 0.2 * r.standard_normal(n)
    P["py"] = 0.2 * r.standard_normal(n)
    P["py"][:2] = [-3.0, 3.0]                 # towards the walls
    P["pz"] = 0.1 * r.standard_normal(n)
    P["w"] = 1e-4
    P["q"] = -1
    P["qm"] = -1
    P["sp"] = 2
    P["alive"] = 1
    P["alive"][9] = 6
    P["zeta"][9] = z[FH_STREAM_N] + 0.5 * (z[FH_STREAM_N + 1] - z[FH_STREAM_N])
    P["alive"][10] = 6
    P["zeta"][10] = z[FH_STREAM_N] - 0.1
    P["z"] = 0.0
    P["id"] = np.arange(n, dtype=np.uint64) + (np.uint64(2) << np.uint64(32))
    return F, P


def deposit_input():
    """200 beam electrons of a small bunch deck: most live in the window, some dead, slow or outside the zeta range"""
    r = np.random.default_rng(18)
    n = 200
    P = np.zeros(n, dtype=g3d_amd.PARTICLE)
    P["x"] = 8 + 2 * r.standard_normal(n)
    P["y"] = 8 + 2 * r.standard_normal(n)
    P["zeta"] = 3.75 + 1.5 * r.standard_normal(n)
    P["zeta"][:3] = [-0.3, 7.6, 7.4]           # below the window, above the last half-cell, inside it
    P["px"] = 0.05 * r.standard_normal(n)
    P["py"] = 0.05 * r.standard_normal(n)
    P["pz"] = 200 * (1 + 0.05 * r.standard_normal(n))
    P["pz"][3] = 2.0
    P["w"] = 1e-3 * (1 + 0.1 * r.standard_normal(n))
    P["q"] = -1
    P["q"][4] = 1
    P["qm"] = P["q"]
    P["alive"] = 1
    P["alive"][5:8] = [0, 2, 3]
    return P


KEY_CASES = [(1, 0, 0, 0), (20240611, 3, 9, 1), (2 ** 64 - 1, 12345, 2047, 1), (7, -1, 5, 0)]
KEY_IDS = np.array([0, 1, (1 << 32) + 5, (200 << 32) + 262143, 2 ** 64 - 1], dtype=np.uint64)
