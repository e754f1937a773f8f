"""GPU parity on the free-hand y grid (conducting walls, one tridiagonal system per k_x line and slice) and on
free-hand zeta tables: the MI355X engine against the reference's own CPU build (oracle/_ref, one thread).

The march of the lines repeats the reference's operations in its order, so it is compared bit for bit; the other
paths sum atomics or transform in another order and are held to 1e-12 relative L2 per slice."""
import numpy as np
import pytest

import g3d_amd
import refpy
from decks import LAMBDA0, particle_parity, slice_rel_l2
from test_gpu_engine import SPEC, compare_fields

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not refpy.available(), reason="oracle/_ref not built")]


@pytest.fixture(scope="module", autouse=True)
def one_thread():
    refpy.set_threads(1)


def fh_plasma_deck(nx=32, nz=48, passes=2, smooth=0.0, a0=1.0, n0=1e-2, sponge=6.0, ions=False):
    """uniform plasma and a Gaussian grid laser on a graded free-hand y grid (0.5/k0 on the axis, 1.5/k0 at the
    walls at +-26/k0)"""
    h = 0.5
    d = g3d_amd.Deck(LAMBDA0)
    d.c.nx = nx; d.c.Lx = 60.0
    d.y_table([0.0, 6.0, 14.0], [0.5, 0.5, 1.5], 26.0)
    d.zeta_uniform(nz, (nz - 1) * h)
    d.set(dz=5.0, z_end=1000.0, plasma_on=1, n0=n0, ppc_x=2, ppc_y=2, move_x=0, sponge=sponge, sponge_y=sponge,
          iterations=passes, smooth_length=smooth)
    d.add_grid_laser(a0=a0, waist=8.0, tau=3.0, zeta_c=(nz - 1) * h * 0.6)
    if ions:
        # nitrogen from the neutral atom, ADK through its levels (NIST energies, eV)
        d.add_ion(14.0067 * 1822.888486, 0, 7, atoms=0.05, mobile=True, ionization=True,
                  ip_eV=[14.53413, 29.60125, 47.4453, 77.4735, 97.8901, 552.06733, 667.04612])
    return d


def fh_bunch_deck(nx=32, nmacro=4096, laser=True):
    """bunch + oblique plane wave in vacuum on a free-hand y grid and a free-hand zeta table"""
    d = g3d_amd.Deck(LAMBDA0)
    d.c.nx = nx; d.c.Lx = 40.0
    d.y_table([0.0, 4.0, 12.0], [0.4, 0.4, 1.2], 20.0)
    d.zeta_table([0.0, 10.0, 20.0, 32.0], [0.6, 0.4, 0.4, 0.7], 32.0)
    Lw = float(d.zeta[-1])
    d.set(dz=20.0, z_end=1000.0, sponge=4.0, sponge_y=4.0, move_x=1,
          charge=11.4014089e-12 * nmacro / 262144, gamma=1000.0, sigx=2.0, sigy=2.0, sigz=4.0, emitx=2e-4, emity=2e-4,
          espread=1e-3, zeta_center=0.5 * Lw, nmacro=nmacro, seed=1)
    if laser:
        d.set(laser_on=1, laser_pol=1, a0=1.0, theta=2e-3, tau=30.0, start_fwhm=0.0, stop_fwhm=0.0)
    return d


def pair(fn, **kw):
    return g3d_amd.Sim(fn(**kw)), refpy.RefSim(fn(**kw))


def random_spectra(sim, names, seed, hermitian_x=False):
    """random planes of the right sizes; hermitian_x: x transforms of real rows (as every field of the free-hand
    grid is)"""
    info = sim.info()
    nx = sim._deck.c.nx
    rng = np.random.default_rng(seed)
    out = {}
    for n in names:
        size = sim._size(n) // 16
        if hermitian_x:
            out[n] = np.fft.fft(rng.standard_normal((size // nx, nx)), axis=1).reshape(-1)
        else:
            out[n] = rng.standard_normal(size) + 1j * rng.standard_normal(size)
    assert info["nk"] == nx * sim._deck.c.ny
    return out


def test_fh_grid_tables():
    """the step tables give the reference's node layout: symmetric nodes, walls at the ends, steps within the
    table's range; the free-hand deck needs no Fourier y"""
    d = fh_plasma_deck()
    y, z = d.ynode, d.zeta
    assert d.c.y_fourier == 0 and d.c.move_y == 0
    assert np.allclose(y, -y[::-1], rtol=0, atol=1e-12) and y[len(y) // 2] == 0
    h = np.diff(y)
    assert h.min() > 0.5 - 1e-12 and h.max() < 1.5 + 1e-12
    b = fh_bunch_deck()
    hz = np.diff(b.zeta)
    assert hz.min() > 0.4 - 1e-12 and hz.max() < 0.7 + 1e-12 and b.zeta[0] == 0


@pytest.mark.parametrize("pivots", [1, 0])
def test_fh_march_random_bit_exact(pivots):
    """Simulation::march(D) on the lines (src/sim_kernels.hpp:213-313): the same operations in the same order,
    so the spectra agree bit for bit, with the pivot table formed once per D (pivots=1) and with the pivots formed
    inside the eliminations (pivots=0, the path of the kernel-level g3d_march)"""
    gpu, ref = pair(fh_plasma_deck, nx=16, nz=24)
    gpu.call("line_pivots", pivots)
    arrs = random_spectra(gpu, SPEC + ["j0x", "j0y", "jkx", "jky", "jkz"], 21)
    for sim in (gpu, ref):
        for k, v in arrs.items():
            sim.set(k, v)
        sim.call("march", 2.5)
    for k in SPEC:
        np.testing.assert_array_equal(gpu.array(k), ref.array(k), err_msg=k)


@pytest.mark.parametrize("pivots", [1, 0])
def test_fh_march_slice_bit_exact(pivots):
    """Simulation::march_slice(n, D) on the lines, the front slice and an inner one, both pivot paths"""
    gpu, ref = pair(fh_plasma_deck, nx=16, nz=24)
    gpu.call("line_pivots", pivots)
    arrs = random_spectra(gpu, SPEC + ["j0x", "j0y", "jkx", "jky", "jkz"], 22)
    for sim in (gpu, ref):
        for k, v in arrs.items():
            sim.set(k, v)
        sim.call("march_slice", 23, 1.5)
        sim.call("march_slice", 11, 1.5)
    for k in SPEC:
        np.testing.assert_array_equal(gpu.array(k), ref.array(k), err_msg=k)


def test_fh_march_after_geometry_change():
    """The pivot table depends on the half-cell lengths and k_x as well as D: after hcg is set anew, a march with
    the same D must use the new geometry (the same spectra as a fresh simulation given that geometry first, and as
    the march without the table)"""
    names = SPEC + ["j0x", "j0y", "jkx", "jky", "jkz"]
    sims = [g3d_amd.Sim(fh_plasma_deck(nx=16, nz=24)) for _ in range(3)]
    arrs = random_spectra(sims[0], names, 24)
    hcg = sims[0].array("hcg") * 1.1
    for k, v in arrs.items():
        sims[0].set(k, v)
    sims[0].call("march", 2.5)                 # the table of the old geometry
    sims[2].call("line_pivots", 0)
    for s in sims:
        s.set("hcg", hcg)
        for k, v in arrs.items():
            s.set(k, v)
        s.call("march", 2.5)
    for k in SPEC:
        np.testing.assert_array_equal(sims[0].array(k), sims[1].array(k), err_msg=k)
        np.testing.assert_array_equal(sims[0].array(k), sims[2].array(k), err_msg=k)


def test_fh_sponge_and_real_fields_random():
    """the bound potential per line (LU factors of k_x^2 - Lap_y), the masks, R_b added back, the real fields
    (src/sim_fields.cpp:156-307)"""
    gpu, ref = pair(fh_plasma_deck, nx=16, nz=24)
    arrs = random_spectra(gpu, SPEC + ["jkz"], 23, hermitian_x=True)
    for sim in (gpu, ref):
        for k, v in arrs.items():
            sim.set(k, v)
        sim.call("sponge_and_real_fields")
    compare_fields(gpu, ref)


def test_fh_stream_spectra_random():
    """x transforms per row, the (1,2,1) filters along y, node 0, the smoothing solves (src/sim_plasma.cpp:
    119-184)"""
    gpu, ref = pair(fh_plasma_deck, nx=16, nz=24, smooth=1.5)
    rng = np.random.default_rng(24)
    J = rng.uniform(-1, 1, gpu._size("jsr") // 8)
    for sim in (gpu, ref):
        sim.set("jsr", J)
        sim.call("stream_spectra")
    a, b = gpu.array("jsn"), ref.array("jsn")
    assert slice_rel_l2(a, b, 3) < 1e-14


def test_fh_plasma_init_and_steps():
    """init + 3 steps of the streaming plasma on the free-hand y grid: no plasma in the wall cells, particles
    leaving through the walls are lost, the lines marched per pass"""
    gpu, ref = pair(fh_plasma_deck)
    gpu.init(); ref.init()
    compare_fields(gpu, ref)
    for _ in range(3):
        gpu.step(); ref.step()
    compare_fields(gpu, ref, groups=("fields", "jk", "js0"))
    a, b = gpu.array("strm"), ref.array("strm")
    assert len(a) == len(b) > 0
    np.testing.assert_array_equal(a["id"], b["id"])
    np.testing.assert_array_equal(a["alive"], b["alive"])
    particle_parity(a, b, fh_plasma_deck())


def test_fh_plasma_smoothing():
    gpu, ref = pair(fh_plasma_deck, smooth=1.5)
    gpu.init(); ref.init()
    gpu.step(); ref.step()
    compare_fields(gpu, ref, groups=("fields", "jk", "js0"))


def test_fh_bunch_init_and_steps():
    """co-moving start field in the eigenvectors of -Lap_y, the relaxation, the beam deposit on nodes and
    half-nodes, 2 steps; free-hand zeta table"""
    gpu, ref = pair(fh_bunch_deck)
    gpu.init(); ref.init()
    pa, pb = gpu.array("parts"), ref.array("parts")
    assert len(pa) == len(pb) > 0
    for f in ("x", "y", "zeta", "px", "py", "pz", "w"):
        np.testing.assert_array_equal(pa[f], pb[f])
    compare_fields(gpu, ref)
    for _ in range(2):
        gpu.step(); ref.step()
    print("after 2 steps:", compare_fields(gpu, ref))
    pa, pb = gpu.array("parts"), ref.array("parts")
    np.testing.assert_array_equal(pa["alive"], pb["alive"])
    particle_parity(pa, pb, fh_bunch_deck())


def canonical(p):
    """the streaming list in an order both codes share: by id (freed electrons carry their ion's id), then by
    zeta (electrons of one ion freed in different half-cells)"""
    return p[np.lexsort((np.round(p["zeta"], 6), p["id"]))]


@pytest.mark.parametrize("freehand", [False, True])
def test_ionization_events(freehand):
    """ADK field ionization of nitrogen from the neutral atom (src/sim_ions.cpp, src/sim_stream.hpp:67-101): the
    charge state of every ion and the number of freed electrons equal the reference's; the fields within 1e-12"""
    if freehand:
        gpu, ref = pair(fh_plasma_deck, a0=2.0, ions=True)
    else:
        from decks import plasma_deck

        def deck():
            d = plasma_deck(a0=2.0)
            d.add_ion(14.0067 * 1822.888486, 0, 7, atoms=0.05, mobile=True, ionization=True,
                      ip_eV=[14.53413, 29.60125, 47.4453, 77.4735, 97.8901, 552.06733, 667.04612])
            return d
        gpu, ref = g3d_amd.Sim(deck()), refpy.RefSim(deck())
    gpu.init(); ref.init()
    for _ in range(2):
        gpu.step(); ref.step()
    a, b = canonical(gpu.array("strm")), canonical(ref.array("strm"))
    assert len(a) == len(b)
    ions = (b["sp"] >= 3) & (b["sp"] < 50)          # ion species; freed electrons are 50 + species
    assert ions.any() and b["q"][ions].sum() > 0, "the deck must ionize"
    assert (a["sp"] >= 50).sum() == (b["sp"] >= 50).sum()
    np.testing.assert_array_equal(a["sp"], b["sp"])
    np.testing.assert_array_equal(a["q"], b["q"])
    np.testing.assert_array_equal(a["alive"], b["alive"])
    compare_fields(gpu, ref, groups=("fields", "jk", "js0"))
