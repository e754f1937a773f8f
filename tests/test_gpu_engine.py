"""GPU parity: the MI355X engine against the reference's own CPU build (oracle/_ref, one thread: deterministic) on
the same decks.  Fields within 1e-12 relative L2 per slice (FP64), particle indices and flags bit-exact.

The reference is run in this process through oracle/refpy.py; its library travels with the tree."""
import numpy as np
import pytest

import g3d_amd
import refpy
from decks import F_rel_l2, bunch_deck, group_rel_l2, particle_parity, plasma_deck, slice_rel_l2

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not refpy.available(), reason="oracle/_ref not built")]
TOL = 1e-12
SPEC = ["Rx", "Ry", "Ez", "Bz", "Lx", "Ly"]


@pytest.fixture(scope="module", autouse=True)
def one_thread():
    refpy.set_threads(1)


def pair(deck_fn, **kw):
    return g3d_amd.Sim(deck_fn(**kw)), refpy.RefSim(deck_fn(**kw))


# F is float32 in the reference too (src/sim.hpp:273): two FP64 results that agree to ~1e-16 still round to
# neighbouring floats now and then, so its bar is the float rounding, not the FP64 tolerance.
F_TOL = 1.2e-7
GROUPS = {"fields": SPEC, "jk": ["jkx", "jky", "jkz"], "js0": ["js0x", "js0y", "js0z"]}


def compare_fields(gpu, ref, tol=TOL, groups=("fields",), F=True):
    """relative L2 per slice over each group of spectra taken together, and of the real fields F"""
    info = gpu.info()
    ns, nk = info["ns"], info["nk"]
    out = {}
    for g in groups:
        names = GROUPS[g]
        planes = ns if g == "fields" else ns + 1
        out[g] = group_rel_l2([gpu.array(n) for n in names], [ref.array(n) for n in names], planes)
    bad = {k: v for k, v in out.items() if not v <= tol}
    if F:
        out["F"] = F_rel_l2(gpu.array("F"), ref.array("F"), ns, nk)
        if not out["F"] <= F_TOL:
            bad["F"] = out["F"]
    assert not bad, "relative L2 per slice above tolerance: %s (all: %s)" % (bad, out)
    return out


def test_march_random_fields():
    """Simulation::march(D) on random spectra (src/sim_fields.cpp:7-39)."""
    gpu, ref = pair(plasma_deck, nx=16, ny=16, nz=24)
    rng = np.random.default_rng(3)
    for sim in (gpu, ref):
        rng = np.random.default_rng(3)
        for n in SPEC + ["j0x", "j0y", "jkx", "jky", "jkz"]:
            a = rng.standard_normal(gpu._size(n) // 16) + 1j * rng.standard_normal(gpu._size(n) // 16)
            sim.set(n, a)
        sim.call("march", 2.5)
    compare_fields(gpu, ref, tol=1e-13)


@pytest.mark.parametrize("nx,ny,nz", [(8, 4, 25), (16, 16, 26), (32, 8, 7)])
def test_march_staged_and_plain_bit_exact(nx, ny, nz, monkeypatch):
    """Simulation::march(D) (src/sim_fields.cpp:7-39) on random spectra, by the staged march (one wave marching 64
    modes from LDS while three waves stage blocks of three slices) and by the plain loop (G3D_MARCH_STAGED=0), on
    grids whose mode count is below 64 (a workgroup partly beyond the modes) and whose slice count is not a multiple
    of the block: both equal the reference bit for bit"""
    got = {}
    for staged in ("1", "0"):
        monkeypatch.setenv("G3D_MARCH_STAGED", staged)
        gpu, ref = pair(plasma_deck, nx=nx, ny=ny, nz=nz)
        for sim in (gpu, ref):
            rng = np.random.default_rng(nz)
            for n in SPEC + ["j0x", "j0y", "jkx", "jky", "jkz"]:
                a = rng.standard_normal(gpu._size(n) // 16) + 1j * rng.standard_normal(gpu._size(n) // 16)
                sim.set(n, a)
            sim.call("march", 2.5)
        got[staged] = {k: gpu.array(k) for k in SPEC}
        for k in SPEC:
            np.testing.assert_array_equal(got[staged][k], ref.array(k), err_msg=f"{k} staged={staged}")
    for k in SPEC:
        np.testing.assert_array_equal(got["1"][k], got["0"][k], err_msg=k)


def test_sponge_and_real_fields_random():
    """Simulation::sponge_and_real_fields() on random Hermitian spectra (src/sim_fields.cpp:156-307)."""
    gpu, ref = pair(plasma_deck, nx=16, ny=16, nz=24)
    ns, nk = 24, 256
    rng = np.random.default_rng(5)
    for n in SPEC + ["jkz"]:
        planes = ns + 1 if n == "jkz" else ns
        x = rng.standard_normal((planes, 16, 16))
        a = np.fft.fft2(x)                              # spectra of real fields (Hermitian)
        for sim in (gpu, ref):
            sim.set(n, a.reshape(-1))
    for sim in (gpu, ref):
        sim.call("sponge_and_real_fields")
    compare_fields(gpu, ref)


def test_plasma_init():
    gpu, ref = pair(plasma_deck)
    gpu.init(); ref.init()
    compare_fields(gpu, ref)


@pytest.mark.parametrize("passes,extrap,layout", [(2, 1, "cm"), (1, 1, "cm"), (2, 0, "cm"), (2, 1, "split")])
def test_plasma_steps(passes, extrap, layout, monkeypatch):
    """init + 3 steps of the streaming-plasma march (src/sim_plasma.cpp:228-470), with the modes column-major and
    one fused column launch per pass (cm), and with the reference's layout and the split column pass (split)"""
    monkeypatch.setenv("G3D_PASS_CM", "1" if layout == "cm" else "0")
    gpu, ref = pair(plasma_deck, passes=passes, extrap=extrap)
    gpu.init(); ref.init()
    for _ in range(3):
        gpu.step(); ref.step()
    compare_fields(gpu, ref, groups=("fields", "jk", "js0"))
    a, b = gpu.array("strm"), ref.array("strm")
    assert len(a) == len(b)
    np.testing.assert_array_equal(a["id"], b["id"])
    np.testing.assert_array_equal(a["alive"], b["alive"])
    for f in ("x", "y", "zeta", "px", "py", "pz", "z"):
        d = np.abs(a[f] - b[f]).max() / max(np.abs(b[f]).max(), 1e-300)
        assert d < 1e-12, (f, d)
    p = np.sqrt(b["px"] ** 2 + b["py"] ** 2 + b["pz"] ** 2)
    dp = np.sqrt((a["px"] - b["px"]) ** 2 + (a["py"] - b["py"]) ** 2 + (a["pz"] - b["pz"]) ** 2)
    assert np.linalg.norm(dp) <= 1e-12 * np.linalg.norm(p)
    rl_a, rl_b = gpu.array("rho_lines"), ref.array("rho_lines")
    assert np.abs(rl_a - rl_b).max() <= 1e-12 * np.abs(rl_b).max()


def test_plasma_smoothing():
    gpu, ref = pair(plasma_deck, smooth=1.5)
    gpu.init(); ref.init()
    gpu.step(); ref.step()
    compare_fields(gpu, ref)


def test_bunch_init_and_steps():
    """config-0 style: bunch + oblique plane wave in vacuum; init (co-moving start field, relaxation), 3 steps.
    The beam deposit sums in the reference's single-thread order, so the start field, which amplifies the last
    bits of the currents by ~1e4, stays within the FP64 bar."""
    gpu, ref = pair(bunch_deck)
    gpu.init(); ref.init()
    pa, pb = gpu.array("parts"), ref.array("parts")
    assert len(pa) == len(pb) > 0
    for f in ("x", "y", "zeta", "px", "py", "pz", "w"):
        np.testing.assert_array_equal(pa[f], pb[f])     # same generator, same draws
    compare_fields(gpu, ref)
    for _ in range(3):
        gpu.step(); ref.step()
    print("after 3 steps:", compare_fields(gpu, ref))
    pa, pb = gpu.array("parts"), ref.array("parts")
    np.testing.assert_array_equal(pa["alive"], pb["alive"])
    particle_parity(pa, pb, bunch_deck())


def test_stream_spectra_random():
    """Simulation::stream_spectra() on a random real current (src/sim_plasma.cpp:94-117)."""
    gpu, ref = pair(plasma_deck, nx=16, ny=16, nz=24, smooth=2.0)
    rng = np.random.default_rng(9)
    J = rng.uniform(-1, 1, 5 * 256)
    for sim in (gpu, ref):
        sim.set("jsr", J)
        sim.call("stream_spectra")
    a, b = gpu.array("jsn"), ref.array("jsn")
    assert slice_rel_l2(a, b, 3) < 1e-13


def test_plasma_strong_laser_tiles():
    """a0 = 4 in a denser plasma: the electrons are blown out by many cells, so deposits leave their tile's LDS
    window and take the global-atomic path beside the slabs; 2 steps against the reference."""
    gpu, ref = pair(plasma_deck, a0=4.0, n0=3e-2)
    gpu.init(); ref.init()
    for _ in range(2):
        gpu.step(); ref.step()
    compare_fields(gpu, ref, tol=1e-11, groups=("fields", "js0"))
    a, b = gpu.array("strm"), ref.array("strm")
    np.testing.assert_array_equal(a["alive"], b["alive"])
    dx = np.abs(((a["x"] - b["x"]) + 30) % 60 - 30).max()      # periodic box of 60/k0
    assert dx < 1e-9


def test_device_lattice_matches_host(monkeypatch):
    """the fresh plasma generated on the device (tiled Fourier-y electron plasma) is the host loader's list, value
    for value and in the same order, ids included"""
    a, b = g3d_amd.Sim(plasma_deck()), g3d_amd.Sim(plasma_deck())
    monkeypatch.setenv("G3D_HOST_LATTICE", "1")
    b.call("load_plasma", 0.0, 5.0)
    monkeypatch.delenv("G3D_HOST_LATTICE")
    a.call("load_plasma", 0.0, 5.0)
    pa, pb = a.array("strm"), b.array("strm")
    assert len(pa) == len(pb) == 32 * 32 * 4
    for f in pa.dtype.names:
        np.testing.assert_array_equal(pa[f], pb[f], err_msg=f)


def test_window_shift_follows_bunch():
    """the transverse window follows an off-centre bunch (src/sim_setup.cpp:221-280): window_lag 1, the bunch moved
    4.6 cells in x and -3.2 cells in y after init, so step 2 shifts the window by (5, -3) cells (phase factors on the
    spectra, the roll of F, the particles moved); X0, Y0, particles and fields after 4 steps against the reference"""
    mk = lambda: bunch_deck().set(window_lag=1)
    gpu, ref = g3d_amd.Sim(mk()), refpy.RefSim(mk())
    gpu.init(); ref.init()
    i0 = ref.info()
    p = ref.array("parts")
    p["x"] += 4.6 * i0["dx"]
    p["y"] -= 3.2 * i0["dy"]
    gpu.set_particles(p); ref.set_particles(p)
    for _ in range(4):
        gpu.step(); ref.step()
    ig, ir = gpu.info(), ref.info()
    assert ir["X0"] == i0["X0"] + 5 * i0["dx"] and ir["Y0"] == i0["Y0"] - 3 * i0["dy"], (ir["X0"], ir["Y0"])
    assert (ig["X0"], ig["Y0"]) == (ir["X0"], ir["Y0"])
    compare_fields(gpu, ref)
    pa, pb = gpu.array("parts"), ref.array("parts")
    np.testing.assert_array_equal(pa["alive"], pb["alive"])
    particle_parity(pa, pb, mk())


def test_windowed_beam_push_equals_plain():
    """the beam push with its particles binned by slice pair and tile and the fields of each bin staged in LDS
    equals the one-lane-per-particle push that gathers from F, bit for bit: the config-0 style bunch in the plane
    wave over 3 steps (particles and fields; the vacuum step is deterministic), and the 2^20-electron witness of the
    laser-wakefield deck at 512 slices pushed twice through the same developed fields (the plasma pass sums with
    FP64 atomics, so two runs of it differ in the last bits; here both pushes start from one state)"""
    from decks import config2_deck
    a, b = g3d_amd.Sim(bunch_deck(nmacro=16384)), g3d_amd.Sim(bunch_deck(nmacro=16384))
    b.call("bpush_windows", 0)
    a.init(); b.init()
    for _ in range(3):
        a.step(); b.step()
    pa, pb = a.array("parts"), b.array("parts")
    assert len(pa) > 4096
    for f in pa.dtype.names:
        np.testing.assert_array_equal(pa[f], pb[f], err_msg=f)
    np.testing.assert_array_equal(a.array("F"), b.array("F"))
    a.close(); b.close()
    a, b = g3d_amd.Sim(config2_deck(0, nz=512)), g3d_amd.Sim(config2_deck(0, nz=512))
    b.call("bpush_windows", 0)
    a.init(); b.init()
    for _ in range(2):
        a.step()
    b.set_particles(a.array("parts"))
    b.set("F", a.array("F"))
    s, dz = a.info()["s"], a._deck.c.dz
    for k in range(2):
        a.call("push", s + k * dz, s + (k + 1) * dz)
        b.call("push", s + k * dz, s + (k + 1) * dz)
    pa, pb = a.array("parts"), b.array("parts")
    assert (pa["alive"] == 1).sum() > 1_000_000
    for f in pa.dtype.names:
        np.testing.assert_array_equal(pa[f], pb[f], err_msg=f)
