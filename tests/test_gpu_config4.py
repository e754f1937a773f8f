"""The inputs of config 4 against the reference's own CPU build (oracle/_ref, one thread): the grid laser whose
amplitude along zeta is a piecewise-linear a0(zeta) (the reference's envelope_function, given to it as the sum of
the hat functions) and the witness whose current profile is piecewise linear (distribution_z = function, profile_z,
sampled by the inverse of its cumulative distribution).  On the config-4 deck at 32 x 32 modes and 64 slices: the
witness drawn at init equals the reference's particle for particle, and after init and two steps the particles and
fields agree at the bars of the FP64 form.

The gradient of the config-4 loss (the mean p_z of the live witness after the steps) in its 32 envelope and 16
profile values by the reverse of the march (Sim.gradient, g3d_sim_gradient): against central differences of the
engine's loss and of the reference's, against the tangent-linear run of the whole march, the same with checkpoints
every step and every second step, the FP32-mixed
form's against the FP64 form's, through torch.autograd (g3d_autograd.WitnessEnergy), and at the configuration's full
size."""
import numpy as np
import pytest

import g3d_amd
import refpy
from decks import config4_deck, config4_params, particle_parity
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


def medium(env=None, prof=None):
    """config 4 on 64 x 64 modes and 256 slices, 16384 witness electrons, two steps"""
    return config4_deck(env=env, prof=prof, n=64, nz=256, nwit=16384, steps=2)


def run_loss(cls, deck, steps):
    s = cls(deck)
    s.init()
    for _ in range(steps):
        s.step()
    p = s.array("parts")
    if hasattr(s, "close"):
        s.close()
    return float(p[(p["alive"] == 1) & (p["sp"] == 0)]["pz"].mean())


def central(cls, make, which, k, h, steps=2):
    env0, prof0 = config4_params()
    val = []
    for sgn in (1, -1):
        env, prof = env0.copy(), prof0.copy()
        (env if which == "env" else prof)[k] += sgn * h
        val.append(run_loss(cls, make(env, prof), steps))
    return (val[0] - val[1]) / (2 * h)


# the bar of the gradient against central differences: 1e-6 absolute plus 1e-4 of the derivative.  The loss is
# smooth only piecewise (a witness particle's current moves to the next half-cell when its zeta crosses a boundary,
# an electron held over to the next step joins the march at a slice, F is float32), so a difference across such an
# event is no derivative: the best of the steps 1e-3 ... 1e-5 is compared
FD_STEPS = (1e-3, 3e-4, 1e-4, 3e-5, 1e-5)


def fd_bar(g):
    return 1e-6 + 1e-4 * abs(g)


def test_config4_gradient_against_differences():
    env0, prof0 = config4_params()
    s = g3d_amd.Sim(medium(env0, prof0))
    loss, genv, gprof = s.gradient(2, every=1)
    p = s.array("parts")                             # the simulation is left at the end of the steps
    assert s.info()["istep"] == 2
    assert abs(float(p[(p["alive"] == 1) & (p["sp"] == 0)]["pz"].mean()) - loss) <= 1e-12 * abs(loss)
    s.close()
    s = g3d_amd.Sim(medium(env0, prof0))
    loss2, genv2, gprof2 = s.gradient(2, every=2)
    s.close()
    assert genv.shape == (32,) and gprof.shape == (16,)
    assert abs(loss - run_loss(g3d_amd.Sim, medium(env0, prof0), 2)) <= 1e-12 * abs(loss)
    np.testing.assert_allclose(genv2, genv, rtol=1e-10, atol=1e-14 * np.abs(genv).max())
    np.testing.assert_allclose(gprof2, gprof, rtol=1e-10, atol=1e-14 * np.abs(gprof).max())
    for which, k in (("env", 8), ("env", 12), ("env", 20), ("prof", 3), ("prof", 7), ("prof", 12)):
        g = (genv if which == "env" else gprof)[k]
        fds = [central(g3d_amd.Sim, medium, which, k, h) for h in FD_STEPS]
        best = min(abs(f - g) for f in fds)
        print("d loss / d %s[%d]: adjoint %.9e, central differences %s, closest %.2e"
              % (which, k, g, " ".join("%.6e" % f for f in fds), best), flush=True)
        assert g != 0 and best <= fd_bar(g), (which, k, g, fds)


def test_config4_tangent_linear():
    """the tangent-linear run of the whole march (Sim.tangent: the tangents carried forward beside the primal state)
    along one envelope value, one profile value and a random direction equals the adjoint gradient's projection: the
    reverse is the transpose of the linearised march to rounding"""
    env0, prof0 = config4_params()
    s = g3d_amd.Sim(medium(env0, prof0))
    loss, genv, gprof = s.gradient(2, every=1)
    s.close()
    rng = np.random.default_rng(7)
    for name, de, dp in (("env[12]", np.eye(32)[12], np.zeros(16)), ("prof[7]", np.zeros(32), np.eye(16)[7]),
                         ("random", rng.standard_normal(32), rng.standard_normal(16))):
        s = g3d_amd.Sim(medium(env0, prof0))
        l2, dl = s.tangent(2, de, dp)
        s.close()
        ref = float(genv @ de + gprof @ dp)
        print("tangent-linear along %s: %.15e, adjoint %.15e" % (name, dl, ref), flush=True)
        assert abs(l2 - loss) <= 1e-12 * abs(loss)
        assert abs(dl - ref) <= 1e-12 * abs(ref) + 1e-15, (name, dl, ref)


def test_config4_mixed_gradient():
    """the gradient of the FP32-mixed form (its forward, the FP64 reverse at its trajectory) against the FP64 form's:
    the two losses agree to 1e-8 and so do their derivatives, to 1e-5 of the largest (central differences of the
    mixed loss resolve no better than 1e-5: its FP32 field pass rounds at 6e-8)"""
    env0, prof0 = config4_params()
    out = []
    for mixed in (0, 1):
        s = g3d_amd.Sim(config4_deck(env=env0, prof=prof0, n=64, nz=256, nwit=16384, steps=2, mixed=mixed))
        out.append(s.gradient(2, every=1))
        s.close()
    (l0, e0, p0), (l1, e1, p1) = out
    de, dp = np.abs(e1 - e0).max() / np.abs(e0).max(), np.abs(p1 - p0).max() / np.abs(p0).max()
    print("mixed against FP64: loss %.2e, gradients %.2e (envelope) %.2e (profile) apart, relative"
          % (abs(l1 - l0) / abs(l0), de, dp), flush=True)
    assert abs(l1 - l0) <= 1e-8 * abs(l0)
    assert de <= 1e-5 and dp <= 1e-5


def test_config4_gradient_against_reference_differences():
    """the adjoint gradient against central differences of the reference's own build on the same deck, at a step
    where the loss is smooth (one envelope value at 1e-3, one profile value at 1e-5)"""
    refpy.set_threads(16)
    env0, prof0 = config4_params()
    s = g3d_amd.Sim(medium(env0, prof0))
    _, genv, gprof = s.gradient(2, every=1)
    s.close()
    for which, k, h in (("env", 8, 1e-3), ("prof", 3, 1e-5)):
        g = (genv if which == "env" else gprof)[k]
        fr = central(refpy.RefSim, medium, which, k, h)
        print("d loss / d %s[%d]: adjoint %.9e, the reference's central difference (%g) %.9e" % (which, k, g, h, fr),
              flush=True)
        assert abs(fr - g) <= fd_bar(g), (which, k, g, fr)


def test_config4_autograd():
    """the loss and its gradient through torch.autograd: backward gives the engine's gradient times the incoming one"""
    import torch
    from g3d_autograd import WitnessEnergy
    env0, prof0 = config4_params()
    env = torch.tensor(env0, dtype=torch.float64, requires_grad=True)
    prof = torch.tensor(prof0, dtype=torch.float64, requires_grad=True)
    out = WitnessEnergy.apply(env, prof, small_deck, 2, 1)
    (3.0 * out).backward()
    s = g3d_amd.Sim(small_deck(env0, prof0))
    loss, genv, gprof = s.gradient(2, every=1)
    s.close()
    assert abs(out.item() - loss) <= 1e-12 * abs(loss)
    np.testing.assert_allclose(env.grad.numpy(), 3.0 * genv, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(prof.grad.numpy(), 3.0 * gprof, rtol=1e-12, atol=1e-15)


def small_deck(env, prof):
    return config4_deck(env=env, prof=prof, n=32, nz=64, nwit=4096, steps=2)


def test_config4_full_gradient():
    """config 4 at its size (128 x 128 modes, 512 slices, 2^18 witness electrons, 20 steps): the gradient in all 48
    values with a checkpoint every 5 steps and with one every 20 agree; the tangent-linear run along a random direction
    equals the gradient's projection (1e-7: two runs' trajectories differ in the last bits, the plasma deposit adding
    with atomics); central differences (1e-4) in two values agree to the loss's own noise at this size (2 %: the steps
    1e-4 ... 1e-6 scatter by that much)"""
    env0, prof0 = config4_params()
    out = []
    for every in (5, 20):
        s = g3d_amd.Sim(config4_deck(env0, prof0))
        out.append(s.gradient(20, every=every))
        s.close()
    (l5, e5, p5), (l20, e20, p20) = out
    print("config 4: loss %.12e, |grad env| %.4e, |grad prof| %.4e" % (l5, np.linalg.norm(e5), np.linalg.norm(p5)),
          flush=True)
    # the plasma deposit adds with atomics, so two runs of the 20 steps differ in the last bits (1e-12 of the loss)
    de, dp = np.abs(e20 - e5).max() / np.abs(e5).max(), np.abs(p20 - p5).max() / np.abs(p5).max()
    print("config 4: checkpoints every 5 and every 20 steps: loss %.2e, gradients %.2e, %.2e apart (relative)"
          % (abs(l20 - l5) / abs(l5), de, dp), flush=True)
    assert abs(l20 - l5) <= 1e-10 * abs(l5)
    assert de <= 1e-6 and dp <= 1e-6
    rng = np.random.default_rng(11)
    de, dp = rng.standard_normal(32), rng.standard_normal(16)
    s = g3d_amd.Sim(config4_deck(env0, prof0))
    _, dl = s.tangent(20, de, dp)
    s.close()
    ref = float(e5 @ de + p5 @ dp)
    print("config 4: tangent-linear along a random direction %.12e, adjoint %.12e" % (dl, ref), flush=True)
    assert abs(dl - ref) <= 1e-7 * abs(ref)
    for which, k in (("env", 5), ("prof", 3)):
        g = (e5 if which == "env" else p5)[k]
        f = central(g3d_amd.Sim, lambda e, p: config4_deck(e, p), which, k, 1e-4, steps=20)
        print("config 4: d loss / d %s[%d]: adjoint %.6e, central difference (1e-4) %.6e" % (which, k, g, f), flush=True)
        assert abs(f - g) <= 0.02 * abs(g), (which, k, g, f)
