"""Config 0 at full size against the reference's own CPU build (oracle/_ref, one thread: deterministic): the vacuum
deck of 64 x 64 modes and 256 slices, the 2^18-electron bunch at gamma 1000 in the tilted plane wave, init and the
configuration's 50 steps.  The beam deposit sums in the reference's single-thread order, so the cells of every
particle and every alive code are the reference's.  The plane wave's exp, cos and sin are the device's, which differ
from glibc's in the last bit now and then: over the 50 RK4 steps those differences reach the momenta at about 1e-13
and the spectra at about 1e-11 relative L2 per slice (3 steps stay within 1e-12, test_gpu_engine.py), so the bars
here are 1e-10 for the spectra and 1e-12 for positions and momenta, with F at its float rounding."""
import numpy as np
import pytest

import g3d_amd
import refpy
from decks import config0_deck, particle_parity
from test_gpu_engine import compare_fields

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not refpy.available(), reason="oracle/_ref not built")]


def test_config0_full_size_50_steps():
    refpy.set_threads(1)
    deck = config0_deck()
    gpu, ref = g3d_amd.Sim(deck), refpy.RefSim(config0_deck())
    gpu.init(); ref.init()
    assert gpu.info()["np"] == ref.info()["np"] == 1 << 18
    for _ in range(50):
        gpu.step(); ref.step()
    ig, ir = gpu.info(), ref.info()
    assert ig["istep"] == ir["istep"] == 50
    assert (ig["X0"], ig["Y0"]) == (ir["X0"], ir["Y0"])
    print("config 0, 50 steps:", compare_fields(gpu, ref, tol=1e-10))
    pa, pb = gpu.array("parts"), ref.array("parts")
    assert len(pa) == len(pb)
    np.testing.assert_array_equal(pa["alive"], pb["alive"])
    particle_parity(pa, pb, deck)
