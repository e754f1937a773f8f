"""The laser-wakefield deck at full size (256 x 256 modes, 1024 slices, 262144 streaming electrons per slice) with
the trailing witness bunch (2^20 particles, sigma_r = 1.5/k0, sigma_zeta = 3/k0, 0.005 n_c peak = 2.40 pC,
gamma = 200), run in the FP64 form and in the FP32-mixed form from the same start: the fields over all slices within
1e-4 relative L2, the witness mean energy gain within 1e-5 relative, its emittance reported against FP64."""
import numpy as np
import pytest

import g3d_amd
from decks import LAMBDA0

pytestmark = pytest.mark.gpu
SPEC = ["Rx", "Ry", "Ez", "Bz", "Lx", "Ly"]


def config2_deck():
    d = g3d_amd.Deck(LAMBDA0)
    d.c.nx = 256; d.c.Lx = 240.0
    d.y_uniform(256, 240.0)
    d.zeta_uniform(1024, 1023 * 0.25)
    d.set(dz=20.0, z_end=2000.0, sponge=24.0, sponge_y=24.0, move_x=0, move_y=0, plasma_on=1, n0=1e-3,
          ppc_x=2, ppc_y=2, smooth_length=0.0, iterations=2)
    d.add_grid_laser(a0=2.0, waist=30.0, tau=15.0, zeta_c=200.0, z_focus=0.0)
    # the witness in the first bucket behind the laser (plasma wavelength 2 pi / sqrt(1e-3) = 199/k0)
    d.set(charge=2.40e-12, gamma=200.0, sigx=1.5, sigy=1.5, sigz=3.0, emitx=1e-3, emity=1e-3, espread=0.0,
          zeta_center=100.0, nmacro=1 << 20, seed=7)
    return d


def moments(p):
    live = p[p["alive"] == 1]
    w = live["w"]
    g = np.sqrt(1 + live["px"] ** 2 + live["py"] ** 2 + live["pz"] ** 2)
    x = live["x"] - np.average(live["x"], weights=w)
    ux = live["px"] - np.average(live["px"], weights=w)
    emit = np.sqrt(np.average(x * x, weights=w) * np.average(ux * ux, weights=w) - np.average(x * ux, weights=w) ** 2)
    return float(np.average(g, weights=w)), float(emit)


def test_config2_mixed_against_fp64():
    sims = {}
    for name, mixed in (("f64", 0), ("mixed", 1)):
        d = config2_deck()
        d.set(mixed=mixed)
        s = g3d_amd.Sim(d)
        s.init()
        for _ in range(2):
            s.step()
        sims[name] = s
    a, b = sims["mixed"], sims["f64"]
    num = sum(float(np.vdot(a.array(n) - b.array(n), a.array(n) - b.array(n)).real) for n in SPEC)
    den = sum(float(np.vdot(b.array(n), b.array(n)).real) for n in SPEC)
    field_err = np.sqrt(num / den)
    g64, e64 = moments(b.array("parts"))
    gmx, emx = moments(a.array("parts"))
    gain_err = abs((gmx - 200.0) - (g64 - 200.0)) / abs(g64 - 200.0)
    print("config 2, 2 steps: fields rel L2 %.3e; witness gain FP64 %.6e mixed %.6e rel err %.3e; emittance FP64 %.6e "
          "mixed %.6e rel err %.3e" % (field_err, g64 - 200, gmx - 200, gain_err, e64, emx, abs(emx - e64) / e64))
    assert abs(g64 - 200.0) > 1e-3, "the witness must gain or lose energy"
    assert field_err <= 1e-4
    assert gain_err <= 1e-5
