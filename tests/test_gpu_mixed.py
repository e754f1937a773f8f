"""The FP32-mixed plasma pass against the FP64 pass of the same engine on the same deck (the FP64 pass is itself
held to the reference's CPU build by test_gpu_engine.py).  Bars of the mixed form: fields within 1e-4 relative L2
per slice, the witness bunch's mean energy gain within 1e-5 relative."""
import numpy as np
import pytest

import g3d_amd
from decks import LAMBDA0, F_rel_l2, group_rel_l2, plasma_deck

pytestmark = pytest.mark.gpu
SPEC = ["Rx", "Ry", "Ez", "Bz", "Lx", "Ly"]
FIELD_TOL = 1e-4
GAIN_TOL = 1e-5


def run_pair(deck_fn, steps, **kw):
    a, b = deck_fn(**kw), deck_fn(**kw)
    b.set(mixed=1)
    s64, smx = g3d_amd.Sim(a), g3d_amd.Sim(b)
    s64.init(); smx.init()
    for _ in range(steps):
        s64.step(); smx.step()
    return s64, smx


def field_errors(s64, smx):
    info = s64.info()
    ns, nk = info["ns"], info["nk"]
    return {"fields": group_rel_l2([smx.array(n) for n in SPEC], [s64.array(n) for n in SPEC], ns),
            "jk": group_rel_l2([smx.array(n) for n in ("jkx", "jky", "jkz")],
                               [s64.array(n) for n in ("jkx", "jky", "jkz")], ns + 1),
            "F": F_rel_l2(smx.array("F"), s64.array("F"), ns, nk)}


def test_mixed_plasma_fields():
    """uniform plasma and a Gaussian grid laser, init + 3 steps, tiled plasma (Fourier y)"""
    s64, smx = run_pair(plasma_deck, 3, a0=2.0)
    err = field_errors(s64, smx)
    print("mixed vs FP64:", err)
    assert all(v <= FIELD_TOL for v in err.values()), err
    a, b = smx.array("strm"), s64.array("strm")
    np.testing.assert_array_equal(a["id"], b["id"])
    p = np.sqrt(b["px"] ** 2 + b["py"] ** 2 + b["pz"] ** 2)
    dp = np.sqrt((a["px"] - b["px"]) ** 2 + (a["py"] - b["py"]) ** 2 + (a["pz"] - b["pz"]) ** 2)
    assert np.linalg.norm(dp) <= FIELD_TOL * np.linalg.norm(p)


def witness_deck(nx=32, ny=32, nz=64):
    """config 2 shrunk: the plasma deck plus a trailing witness bunch at gamma 200"""
    d = plasma_deck(nx=nx, ny=ny, nz=nz, a0=2.0)
    Lw = (nz - 1) * 0.5
    d.set(charge=0.531 * 1e-12 * 4096 / 1048576, gamma=200.0, sigx=1.5, sigy=1.5, sigz=3.0, emitx=1e-3, emity=1e-3,
          espread=0.0, zeta_center=0.25 * Lw, nmacro=4096, seed=7)
    return d


def test_mixed_witness_energy_gain():
    """plasma + witness bunch, 3 steps: the witness mean energy gain of the mixed run against the FP64 run"""
    s64, smx = run_pair(witness_deck, 3)
    err = field_errors(s64, smx)
    g = {}
    for name, s in (("f64", s64), ("mixed", smx)):
        p = s.array("parts")
        live = p[p["alive"] == 1]
        g[name] = float(np.average(np.sqrt(1 + live["px"] ** 2 + live["py"] ** 2 + live["pz"] ** 2), weights=live["w"]))
    gain64, gainmx = g["f64"] - 200.0, g["mixed"] - 200.0
    rel = abs(gainmx - gain64) / abs(gain64)
    print("mixed vs FP64:", err, "gain", gain64, gainmx, "rel", rel)
    assert abs(gain64) > 1e-3, "the witness must gain or lose energy"
    assert rel <= GAIN_TOL, rel
    assert err["fields"] <= FIELD_TOL, err
