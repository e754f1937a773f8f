"""The inputs of config 4 against the reference's own CPU build (oracle/_ref, one thread): the grid laser whose
amplitude along zeta is a piecewise-linear a0(zeta) (the reference's envelope_function, given to it as the sum of
the hat functions) and the witness whose current profile is piecewise linear (distribution_z = function, profile_z,
sampled by the inverse of its cumulative distribution).  On the config-4 deck at 32 x 32 modes and 64 slices: the
witness drawn at init equals the reference's particle for particle, and after init and two steps the particles and
fields agree at the bars of the FP64 form."""
import numpy as np
import pytest

import g3d_amd
import refpy
from decks import config4_deck, particle_parity
from test_gpu_engine import compare_fields

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not refpy.available(), reason="oracle/_ref not built")]


def small():
    return config4_deck(n=32, nz=64, nwit=4096, steps=2)


def test_config4_inputs_against_reference():
    refpy.set_threads(1)
    gpu, ref = g3d_amd.Sim(small()), refpy.RefSim(small())
    gpu.init(); ref.init()
    a, b = gpu.array("parts"), ref.array("parts")
    assert len(b) > 3000
    for f in ("x", "y", "zeta", "px", "py", "pz", "w"):
        np.testing.assert_array_equal(a[f], b[f], err_msg=f)
    c = small().c
    assert b["zeta"].min() >= c.prof_zmin - 1e-9 and b["zeta"].max() <= c.prof_zmax + 1e-9
    compare_fields(gpu, ref, groups=("fields",))
    for _ in range(2):
        gpu.step(); ref.step()
    a, b = gpu.array("parts"), ref.array("parts")
    live = b["alive"] == 1
    np.testing.assert_array_equal(a["alive"], b["alive"])
    particle_parity(a[live], b[live], small())
    compare_fields(gpu, ref, groups=("fields", "jk", "js0"))


def loss(sim_cls, env, prof, steps=2):
    """the config-4 loss, the witness mean p_z after the steps, of a simulation class on the small deck"""
    s = sim_cls(config4_deck(env=env, prof=prof, n=32, nz=64, nwit=4096, steps=steps))
    s.init()
    for _ in range(steps):
        s.step()
    p = s.array("parts")
    live = p[(p["alive"] == 1) & (p["sp"] == 0)]
    return float(live["pz"].mean())


def test_config4_fd_derivatives_against_reference():
    """central differences (step 1e-6) of the loss in one laser amplitude and one profile value, the engine's against
    the reference's: the same derivative (the forward of config 4 agrees to first order in its 48 parameters)"""
    from decks import config4_params
    refpy.set_threads(1)
    env0, prof0 = config4_params()
    eps = 1e-6
    for which, k in (("env", 12), ("prof", 7)):
        out = {}
        for name, cls in (("gpu", g3d_amd.Sim), ("ref", refpy.RefSim)):
            val = []
            for sgn in (1, -1):
                env, prof = env0.copy(), prof0.copy()
                (env if which == "env" else prof)[k] += sgn * eps
                val.append(loss(cls, env, prof))
            out[name] = (val[0] - val[1]) / (2 * eps)
        print("d loss / d %s[%d]: engine %.9e reference %.9e" % (which, k, out["gpu"], out["ref"]), flush=True)
        assert out["ref"] != 0
        assert abs(out["gpu"] - out["ref"]) <= 1e-5 * abs(out["ref"]) + 1e-7, (which, k, out)
