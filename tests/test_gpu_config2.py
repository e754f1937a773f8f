"""Config 2 at full size: the laser-wakefield deck (256 x 256 modes, 1024 slices, 262144 streaming electrons per slice)
with the trailing witness bunch (2^20 electrons at gamma 200), run in the FP64 form and in the FP32-mixed form from the
same start for the 100 steps of the configuration.  Bars of the mixed form: the real fields F of every slice within
1e-4 relative L2 of the FP64 fields (E and B of the slice together, no slice left out), the witness mean energy gain
within 1e-5 relative; the witness emittance error (the reference's centred moments) is reported."""
import numpy as np
import pytest

import g3d_amd
from decks import F_slice_errors, bunch_moments, config2_deck

pytestmark = pytest.mark.gpu
STEPS = 100
FIELD_TOL = 1e-4
GAIN_TOL = 1e-5


def test_config2_mixed_against_fp64():
    s64, smx = g3d_amd.Sim(config2_deck(0)), g3d_amd.Sim(config2_deck(1))
    s64.init(); smx.init()
    info = s64.info()
    ns, nk = info["ns"], info["nk"]
    worst = []
    for k in range(1, STEPS + 1):
        s64.step(); smx.step()
        if k in (25, 50, STEPS):
            err, nb = F_slice_errors(smx.array("F"), s64.array("F"), ns, nk)
            worst.append((k, float(err.max()), int(err.argmax()), float(nb[err.argmax()] / nb.max())))
            print("config 2, step %d: F per-slice relative L2 max %.3e at slice %d (norm %.2e of the largest), "
                  "median %.3e" % (k, err.max(), err.argmax(), nb[err.argmax()] / nb.max(), np.median(err)),
                  flush=True)
    g0 = 200.0
    _, g64, ex64, ey64 = bunch_moments(s64.array("parts"))
    _, gmx, exmx, eymx = bunch_moments(smx.array("parts"))
    gain_err = abs((gmx - g0) - (g64 - g0)) / abs(g64 - g0)
    print("config 2, %d steps: witness gain FP64 %.9e mixed %.9e rel err %.3e; eps_nx FP64 %.6e rel err %.3e; "
          "eps_ny rel err %.3e" % (STEPS, g64 - g0, gmx - g0, gain_err, ex64, abs(exmx - ex64) / ex64,
                                   abs(eymx - ey64) / ey64))
    assert abs(g64 - g0) > 1e-2, "the witness must gain or lose energy"
    for k, e, n, r in worst:
        assert e <= FIELD_TOL, (k, e, n, r)
    assert gain_err <= GAIN_TOL, gain_err
