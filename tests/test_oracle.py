"""CPU tests of the oracle (oracle/g3d_oracle.c): pinned against the reference's own build (the golden vectors of
tests/golden/, written by oracle/_ref) and against what the mathematics fixes (direct DFT, spline moments, Boris
invariants, the Landau rate of hydrogen, the published SplitMix64 vectors).  No GPU."""
import math
import os

import numpy as np
import pytest

import g3d_amd
import golden_inputs as gi
import oraclepy as orc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
pytestmark = pytest.mark.skipif(not orc.available(), reason="oracle not built (make -C oracle oracle)")


def gold(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("nx,ny", [(2, 2), (8, 16), (32, 4), (64, 64)])
def test_fft_matches_direct_dft(nx, ny):
    """forward sign -1, inverse scaled by 1/(nx ny) (src/fftmany.hpp:2-3); brute force DFT on small sizes"""
    r = np.random.default_rng(nx * ny)
    x = r.standard_normal((2, ny, nx)) + 1j * r.standard_normal((2, ny, nx))
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    direct = np.zeros_like(x)
    for ky in range(ny):
        for kx in range(nx):
            direct[:, ky, kx] = (x * np.exp(-2j * np.pi * (kx * ii / nx + ky * jj / ny))).sum(axis=(1, 2))
    f = orc.fft_many(nx, ny, True, x, 0)
    assert np.abs(f - direct).max() <= 1e-12 * np.abs(direct).max()
    b = orc.fft_many(nx, ny, True, f, 1)
    assert np.abs(b - x).max() <= 1e-14 * np.abs(x).max()


def test_fft_golden():
    g = gold("fft")
    x = g["x"]
    assert np.array_equal(orc.fft_many(8, 16, True, x, 0), g["fwd2d"])    # same butterflies: bit for bit
    assert np.array_equal(orc.fft_many(8, 16, True, x, 1), g["inv2d"])
    assert np.array_equal(orc.fft_many(8, 16, False, x, 0), g["fwd1d"])
    assert np.array_equal(orc.fft_many(8, 16, False, x, 1), g["inv1d"])


# ---------------------------------------------------------------- grids and pushers
def test_tsc_moments():
    """quadratic spline: weights sum to 1, reproduce the position, variance 1/4"""
    import ctypes as C
    for xi in np.linspace(-3.2, 5.7, 37):
        idx = np.zeros(3, dtype=np.int32)
        w = np.zeros(3)
        orc.lib().or_tsc_index(float(xi), idx.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p))
        assert abs(w.sum() - 1) < 1e-15
        assert abs((w * idx).sum() - xi) < 1e-13
        assert abs((w * (idx - xi) ** 2).sum() - 0.25) < 1e-13
        assert idx[1] == math.floor(xi + 0.5) and np.all(np.diff(idx) == 1)


def test_boris_invariants():
    import ctypes as C
    r = np.random.default_rng(3)
    for _ in range(20):
        u = r.standard_normal(3) * 5
        B = r.standard_normal(3)
        u2 = u.copy()
        orc.lib().or_boris(u2.ctypes.data_as(C.c_void_p), np.zeros(3).ctypes.data_as(C.c_void_p),
                           B.ctypes.data_as(C.c_void_p), -1.0, 0.7)
        assert abs(np.linalg.norm(u2) - np.linalg.norm(u)) < 1e-13 * np.linalg.norm(u)   # pure rotation
        E = r.standard_normal(3)
        u3 = u.copy()
        orc.lib().or_boris(u3.ctypes.data_as(C.c_void_p), E.ctypes.data_as(C.c_void_p),
                           np.zeros(3).ctypes.data_as(C.c_void_p), -1.0, 0.7)
        assert np.allclose(u3, u - 0.7 * E, rtol=0, atol=1e-14 * (1 + np.abs(u).max()))   # pure kick q dt E


def test_push_golden():
    """RK4 in z with the gathers and the oblique plane wave (src/sim_kernels.hpp:98-149): bit for bit"""
    g = gold("push")
    deck = gi.bunch_small_deck()
    grid = orc.Grid(deck)
    out = orc.bunch_push(grid, deck, g["F"], g["parts_in"], 0.0, 20.0, float(g["X0"]), float(g["Y0"]))
    ref = g["parts_out"]
    for f in ("x", "y", "zeta", "px", "py", "pz", "z", "alive"):
        assert np.array_equal(out[f], ref[f]), f


def test_stream_pass_golden():
    """one streaming half-cell pass (src/sim_stream.hpp:104-171), in list order: bit for bit, deposit included"""
    g = gold("stream_pass")
    deck = gi.small_deck()
    grid = orc.Grid(deck)
    p, J, _, _ = orc.stream_halfcell(grid, deck, g["F"], g["parts_in"], gi.STREAM_N, 0.0, 5.0, float(g["X0"]), float(g["Y0"]))
    ref = g["parts_out"]
    for f in ("x", "y", "zeta", "px", "py", "pz", "z", "alive"):
        assert np.array_equal(p[f], ref[f]), f
    assert (p["alive"] == 5).sum() >= 1                     # the fast electrons were trapped
    assert np.array_equal(J, g["jsr"])


# ---------------------------------------------------------------- field solver pieces
def test_march_golden():
    g = gold("march")
    deck = gi.small_deck()
    grid = orc.Grid(deck)
    arr = {k[3:]: g[k] for k in g.files if k.startswith("in_")}
    out = orc.march(grid, arr, gi.MARCH_D)
    for k in ("Rx", "Ry", "Ez", "Bz", "Lx", "Ly"):
        assert np.array_equal(out[k], g["out_" + k]), k


def test_march_freehand_golden():
    """the march of the k_x lines between conducting walls (free-hand y) against the reference's own output"""
    g = gold("march_fh")
    deck = gi.small_fh_deck()
    assert np.array_equal(deck.ynode, g["ynode"])
    grid = orc.Grid(deck)
    arr = {k[3:]: g[k] for k in g.files if k.startswith("in_")}
    out = orc.march(grid, arr, gi.MARCH_D)
    for k in ("Rx", "Ry", "Ez", "Bz", "Lx", "Ly"):
        assert np.array_equal(out[k], g["out_" + k]), k


def test_march_freehand_walls_and_front():
    """the wall rules of the line march: E_z, R_x, L_x vanish on the wall node 0 of every slice below the front,
    and the front slice keeps only R (L = E_z = B_z = 0)"""
    deck = gi.small_fh_deck()
    grid = orc.Grid(deck)
    r = np.random.default_rng(5)
    N, nk, nx = grid.N, grid.nk, grid.nx
    arr = {k: r.standard_normal(N * nk) + 1j * r.standard_normal(N * nk) for k in ("Rx", "Ry", "Ez", "Bz", "Lx", "Ly")}
    arr.update({k: r.standard_normal((N + 1) * nk) + 1j * r.standard_normal((N + 1) * nk)
                for k in ("j0x", "j0y", "jkx", "jky", "jkz")})
    out = orc.march(grid, arr, 1.0)
    node0 = np.zeros((N, deck.c.ny, nx), bool)
    node0[:, 0, :] = True
    for k in ("Ez", "Rx", "Lx"):
        assert np.all(out[k].reshape(N, deck.c.ny, nx)[:-1, 0, :] == 0), k
    front = slice((N - 1) * nk, N * nk)
    for k in ("Ez", "Bz", "Lx", "Ly"):
        assert np.all(out[k][front] == 0), k


def test_march_vacuum_free_wave():
    """a mode with no current and D = 0 (the level-s0 solve) keeps R and gives L = O(h k) (src/sim_kernels.hpp:184)"""
    deck = gi.small_deck()
    grid = orc.Grid(deck)
    N, nk = grid.N, grid.nk
    arr = {k: np.zeros(N * nk, complex) for k in ("Rx", "Ry", "Ez", "Bz", "Lx", "Ly")}
    arr.update({k: np.zeros((N + 1) * nk, complex) for k in ("j0x", "j0y", "jkx", "jky", "jkz")})
    arr["Rx"][:] = 1.0
    out = orc.march(grid, arr, 0.0)
    assert np.array_equal(out["Rx"], arr["Rx"])                 # D = 0: R is kept
    k0 = slice(0, N * nk, nk)                                   # the k = 0 mode: no coupling at all
    assert np.all(out["Ez"][k0] == 0) and np.all(out["Lx"][k0] == 0)


def test_stream_spectra_golden():
    g = gold("stream_spectra")
    deck = gi.small_deck(smooth=1.5)
    grid = orc.Grid(deck)
    jsn = orc.stream_spectra(grid, g["J"], float(g["a2"]), 1)
    assert np.array_equal(jsn, g["jsn"])


# ---------------------------------------------------------------- ADK and the draws
def test_adk_landau_hydrogen():
    """H-like level at the Rydberg energy: the rate is the Landau rate 4 w_a (E_a/E) exp(-2 E_a/3E) (SI units)"""
    Ea, wa, UH = 5.14220674763e11, 4.1341373335e16, 13.605693122994
    lev = orc.adk_levels([UH], 0, 1, 1.0, 1.0, 1e-15)[0]
    for E in (1e10, 3e10, 1e11, 2e11):
        ref = 4 * wa * (Ea / E) * math.exp(-2 * Ea / (3 * E))
        assert abs(orc.adk_rate(lev, E) - ref) <= 1e-12 * ref


def test_adk_golden():
    g = gold("adk")
    E0, omega0 = gi.units()
    table = orc.adk_levels(g["ip"], 0, 7, E0, omega0, 2 * 500.0)
    assert np.array_equal(table, g["table"])
    # the rate grows with the field below the peak, and is 0 below the skip threshold
    for L in table:
        assert orc.adk_rate(L, 0.5 * L[4]) == 0.0
        assert 0 < orc.adk_rate(L, L[4] * 1.01) < orc.adk_rate(L, min(L[3], 2 * L[4]))


def test_mix64_published_vectors():
    """SplitMix64 (Steele, Lea, Flood 2014) seeded with 1234567: the published first five outputs"""
    g_ = 0x9E3779B97F4A7C15
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
            16408922859458223821]
    assert [orc.mix64((1234567 + k * g_) % 2 ** 64) for k in range(5)] == want
    g = gold("mix64")
    assert [orc.mix64(int(k)) for k in g["keys"]] == [int(v) for v in g["mix"]]
    assert [orc.lib().or_uniform01(int(k)) for k in g["keys"]] == list(g["u01"])


def test_real_fields_matches_numpy():
    """E_x = (R_x + L_x)/2 ... in real space from the packed inverse transforms (src/sim_fields.cpp:270-293)"""
    grid = orc.Grid(gi.small_deck())
    r = np.random.default_rng(4)
    arr = {k: np.fft.fft2(r.standard_normal((3, 8, 8))).reshape(-1) for k in ("Rx", "Ry", "Lx", "Ly", "Ez", "Bz")}
    F = orc.real_fields(grid, arr, 3).reshape(6, 5, 64)[:, :3]
    re = {k: np.fft.ifft2(arr[k].reshape(3, 8, 8)).real.reshape(3, 64) for k in arr}
    want = [0.5 * (re["Rx"] + re["Lx"]), 0.5 * (re["Ry"] + re["Ly"]), re["Ez"], 0.5 * (re["Ly"] - re["Ry"]),
            0.5 * (re["Rx"] - re["Lx"]), re["Bz"]]
    for c in range(6):
        assert np.abs(F[c] - want[c]).max() <= 1e-6 * np.abs(want[c]).max()


# ---------------------------------------------------------------- the paths pinned by the later goldens
def test_stream_ions_golden():
    """one streaming pass with ADK nitrogen, mobile and fixed ions (src/sim_stream.hpp:67-171): charge states, the
    freed electrons in list order and their count, and the deposit bit for bit against the reference"""
    g = gold("stream_ions")
    deck = gi.ion_deck()
    grid = orc.Grid(deck)
    seed, step, n, pas = gi.ION_KEY
    key = orc.lib().or_key_slice(orc.lib().or_key_step(seed, step), n, pas)
    assert key == int(g["key"])
    cap = int(g["cap"])
    p, J, born, cnt = orc.stream_halfcell(grid, deck, g["F"], g["parts_in"], n, 0.0, 5.0, float(g["X0"]),
                                          float(g["Y0"]), key=key, bcap=cap)
    ref = g["parts_out"]
    for f in ("x", "y", "zeta", "px", "py", "pz", "z", "q", "qm", "alive", "sp"):
        assert np.array_equal(p[f], ref[f]), f
    assert np.array_equal(J, g["jsr"])
    assert cnt == len(g["born"]) and np.array_equal(born, g["born"])
    assert cnt >= 10 and (p["q"] != g["parts_in"]["q"]).sum() >= 10        # the deck ionizes
    # every freed electron is the clone of an ion (same id and position, q = qm = -1, species 50 + species index)
    ions = {int(i): s for i, s in zip(p["id"], p["sp"]) if 3 <= s < 50}
    assert all(b["q"] == -1 and b["qm"] == -1 and b["sp"] == 50 + ions[int(b["id"])] - 3 for b in born)
    # charge is conserved: the charge gained by the ions equals the electrons freed
    assert (p["q"] - g["parts_in"]["q"])[p["sp"] >= 3].sum() == cnt


def test_stream_freehand_golden():
    """one streaming pass on a free-hand y grid between walls and a free-hand zeta table (the table branches of
    Axis::locate, src/grid.hpp:35-46; the node and half-node deposit, the wall loss): bit for bit"""
    g = gold("stream_fh")
    deck = gi.fh_stream_deck()
    assert np.array_equal(deck.ynode, g["ynode"]) and np.array_equal(deck.zeta, g["zeta"])
    grid = orc.Grid(deck)
    assert not grid.uniform and not grid.y_uniform
    p, J, _, _ = orc.stream_halfcell(grid, deck, g["F"], g["parts_in"], gi.FH_STREAM_N, 0.0, 5.0, float(g["X0"]),
                                     float(g["Y0"]))
    ref = g["parts_out"]
    for f in ("x", "y", "zeta", "px", "py", "pz", "z", "alive"):
        assert np.array_equal(p[f], ref[f]), f
    assert (p["alive"] == 0).sum() >= 2                     # lost through the walls
    assert np.array_equal(J, g["jsr"])


def test_axis_locate_table_brute_force():
    """the table locate (src/grid.hpp:35-46) against a linear scan: X[i] <= x < X[i+1] inside, the end cells
    continued outside, f = (x - X[i]) / (X[i+1] - X[i])"""
    import ctypes as C
    X = np.array(gi.fh_stream_deck().zeta)
    ax = orc.Axis(orc.dptr(X), len(X), 0, 0.0, 1.0)
    i, f = C.c_int(), C.c_double()
    r = np.random.default_rng(9)
    for x in np.concatenate([r.uniform(X[0] - 2, X[-1] + 2, 300), X, X[:-1] + 1e-13]):
        orc.lib().or_axis_locate(C.byref(ax), float(x), C.byref(i), C.byref(f))
        if X[0] <= x < X[-1]:
            k = max(j for j in range(len(X) - 1) if X[j] <= x)
            assert i.value == k and abs(f.value - (x - X[k]) / (X[k + 1] - X[k])) < 1e-13
        else:
            lo = 0 if x < X[0] else len(X) - 2
            t = (x - X[lo if x < X[0] else -1]) / (X[lo + 1] - X[lo])
            assert i.value + f.value == pytest.approx((0 if x < X[0] else len(X) - 1) + t, abs=1e-12)


def test_bunch_deposit_golden():
    """the beam deposit of currents() (src/sim_particles.cpp:56-107): jw bit for bit in list order, and the charge
    of the deposited particles recovered from j_z (sum j_z h dx dy = sum q w over the particles in the window)"""
    g = gold("bunch_deposit")
    deck = gi.bunch_small_deck()
    grid = orc.Grid(deck)
    J = orc.bunch_deposit(grid, g["parts"])
    assert np.array_equal(J, g["jw"])
    P = g["parts"]
    N, nk = grid.N, grid.nk
    jz = J.reshape(3, N + 3, nk)[2]
    live = (P["alive"] == 1) & (P["zeta"] >= 0) & (P["zeta"] < grid.zeta[-1])
    h = grid.hcg[2]
    assert abs(jz.sum() * h * grid.dx * grid.dy - (P["q"] * P["w"])[live].sum()) <= 1e-12 * np.abs(P["w"]).sum()
    assert live.sum() < len(P)                              # dead and outside particles are left out


def test_key_chain_golden():
    """ion_key_step / ion_key_slice / ion_key_particle (src/ionization.hpp:44-46) against the reference, including
    the wrap of the seed product and a negative step"""
    g = gold("keys")
    for c in range(len(g["seed"])):
        ks, kl, kp = orc.key_chain(int(g["seed"][c]), int(g["step"][c]), int(g["n"][c]), int(g["pas"][c]),
                                   g["ids"])
        assert ks == int(g["key_step"][c]) and kl == int(g["key_slice"][c])
        assert kp == [int(v) for v in g["key_particle"][c]]
    # the keys of neighbouring particles, slices and steps differ (a fixed point would repeat every draw)
    flat = g["key_particle"].reshape(-1)
    assert len(set(int(v) for v in flat)) == flat.size
