"""Config 3 at its full transverse size against the reference's own CPU build (oracle/_ref): 512 Fourier x modes
by 386 free-hand y cells between conducting walls at +-200/k0 (0.2/k0 on the axis, growing to 2/k0), electrons at
2 x 1 per cell, mobile pre-ionized hydrogen, nitrogen from the neutral atom with ADK ionization through its seven
levels, the a0 = 2.5 Gaussian grid laser (waist 25/k0, tau 12/k0).  Along zeta a 38-slice table (0.1 -> 0.5/k0)
stands in for the 2048 slices, so that the reference finishes a step in about half a minute; the transverse work
per half-cell is the full configuration's (1.18 M streaming particles, plus the freed electrons).

One step, then: the charge state of every ion, the freed-electron count and every particle's cells equal the
reference's, positions and momenta within 1e-12 relative L2, the fields within 1e-12 per slice.

The FP32-mixed form of config 3 at its full size (2048 slices on the 0.1 -> 0.5/k0 table over 615.5/k0) against the
FP64 form from the same start: the real fields of every slice within 1e-4 relative L2 after each step; the freed
electrons and the nitrogen charge-state populations are reported and stay within 1 % of the FP64 form's."""
import numpy as np
import pytest

import g3d_amd
import refpy
from decks import LAMBDA0, F_slice_errors, particle_parity
from test_gpu_engine import compare_fields
from test_gpu_freehand import canonical

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not refpy.available(), reason="oracle/_ref not built")]

# NIST ionization energies of nitrogen [eV] (the reference's atomic-data row N, src/atomic_data.cpp:36)
N_IP = [14.53413, 29.60125, 47.4453, 77.4735, 97.8901, 552.06733, 667.04612]
AMU = 1822.888486


def config3_deck():
    Lw = 8.0
    d = g3d_amd.Deck(LAMBDA0)
    d.c.nx = 512; d.c.Lx = 400.0
    d.y_table([0.0, 18.4, 30.4], [0.2, 0.2, 2.0], 200.0)
    d.zeta_table([0.0, 2.0, 4.0], [0.1, 0.1, 0.5], Lw)
    d.set(dz=5.0, z_end=1000.0, plasma_on=1, n0=1e-3, ppc_x=2, ppc_y=1, move_x=0, sponge=8.0, sponge_y=8.0,
          iterations=2)
    d.add_grid_laser(a0=2.5, waist=25.0, tau=12.0, zeta_c=0.6 * Lw)
    d.add_ion(1.00794 * AMU, 1, 1, atoms=1.0, mobile=True, ionization=False)
    d.add_ion(14.0067 * AMU, 0, 7, atoms=0.01, mobile=True, ionization=True, ip_eV=N_IP)
    return d


def test_config3_transverse_full_size():
    deck = config3_deck()
    assert deck.c.nx == 512 and deck.c.ny == 386 and deck.c.nzeta == 38
    refpy.set_threads(16)
    gpu, ref = g3d_amd.Sim(deck), refpy.RefSim(config3_deck())
    gpu.init(); ref.init()
    gpu.step(); ref.step()
    a, b = canonical(gpu.array("strm")), canonical(ref.array("strm"))
    print("streaming list: %d (freed electrons %d)" % (len(b), (b["sp"] >= 50).sum()))
    assert len(a) == len(b) > 1_100_000
    assert (b["sp"] >= 50).sum() > 0, "the deck must ionize"
    np.testing.assert_array_equal(a["sp"], b["sp"])
    np.testing.assert_array_equal(a["q"], b["q"])
    np.testing.assert_array_equal(a["alive"], b["alive"])
    live = b["alive"] == 1
    particle_parity(a[live], b[live], deck)
    compare_fields(gpu, ref, groups=("fields", "jk", "js0"))


def config3_full_deck(mixed):
    """config 3 at full size: the transverse grid above, 2048 slices from a 6-point table (0.1 ... 0.5/k0) over
    615.5/k0, the laser 40/k0 behind the front of the window"""
    d = g3d_amd.Deck(LAMBDA0)
    d.c.nx = 512; d.c.Lx = 400.0
    d.y_table([0.0, 18.4, 30.4], [0.2, 0.2, 2.0], 200.0)
    d.zeta_table([0.0, 20.0, 60.0, 100.0, 200.0, 300.0], [0.1, 0.1, 0.15, 0.2, 0.3, 0.5], 615.5)
    d.set(dz=5.0, z_end=1000.0, plasma_on=1, n0=1e-3, ppc_x=2, ppc_y=1, move_x=0, sponge=8.0, sponge_y=8.0,
          iterations=2, mixed=mixed)
    d.add_grid_laser(a0=2.5, waist=25.0, tau=12.0, zeta_c=float(d.zeta[-1]) - 40.0)
    d.add_ion(1.00794 * AMU, 1, 1, atoms=1.0, mobile=True, ionization=False)
    d.add_ion(14.0067 * AMU, 0, 7, atoms=0.01, mobile=True, ionization=True, ip_eV=N_IP)
    return d


def test_config3_mixed_against_fp64():
    s64, smx = g3d_amd.Sim(config3_full_deck(0)), g3d_amd.Sim(config3_full_deck(1))
    s64.init(); smx.init()
    info = s64.info()
    assert info["N"] == 2048 and smx.info()["N"] == 2048
    ns, nk = info["ns"], info["nk"]
    for k in range(1, 4):
        s64.step(); smx.step()
        err, nb = F_slice_errors(smx.array("F"), s64.array("F"), ns, nk)
        h64, hmx = s64.array("ion_stats1"), smx.array("ion_stats1")
        f64, fmx = s64.info()["nstrm"], smx.info()["nstrm"]
        print("config 3 mixed, step %d: F per-slice relative L2 max %.3e at slice %d (norm %.2e of the largest), "
              "median %.3e; streaming list FP64 %d mixed %d; N charge states FP64 %s mixed %s"
              % (k, err.max(), err.argmax(), nb[err.argmax()] / nb.max(), np.median(err), f64, fmx,
                 np.array2string(h64, precision=4), np.array2string(hmx, precision=4)), flush=True)
        assert err.max() <= 1e-4, (k, err.max(), int(err.argmax()))
        assert abs(fmx - f64) <= 0.01 * f64
        np.testing.assert_allclose(hmx, h64, rtol=1e-2, atol=1e-2 * np.abs(h64).max())
    assert h64[-1] > 0, "the deck must ionize"
