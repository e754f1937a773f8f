"""Outputs and checkpoints of the engine against the reference's own (its CPU build, oracle/_ref, one thread): from
the same state (the reference's particles, fields and density lines copied into the engine) bunches.txt,
slices_NNNNNN.txt, the particle dump, the frame / fields_NNNNNN.bin file and the axis lineouts are the same text and
the same bytes, and diag.txt the same numbers; a G3DCKPT1 checkpoint written by either is read by the other with
every array intact, and a run restarted from a checkpoint continues as the uninterrupted run."""
import numpy as np
import pytest

import g3d_amd
import refpy
from decks import F_rel_l2, bunch_deck, group_rel_l2, plasma_deck

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not refpy.available(), reason="oracle/_ref not built")]
SPEC = ["Rx", "Ry", "Ez", "Bz", "Lx", "Ly"]
CKPT = SPEC + ["j0x", "j0y"]


@pytest.fixture(scope="module", autouse=True)
def one_thread():
    refpy.set_threads(1)


def same_state(gpu, ref, names=("F",)):
    """copy the reference's particles and arrays into the engine"""
    gpu.set_particles(ref.array("parts"))
    for n in names:
        gpu.set(n, ref.array(n))


def test_bunch_moments_and_dumps(tmp_path):
    gpu, ref = g3d_amd.Sim(bunch_deck()), refpy.RefSim(bunch_deck())
    gpu.init(); ref.init()
    for _ in range(2):
        gpu.step(); ref.step()
    same_state(gpu, ref)
    for sim, tag in ((gpu, "g"), (ref, "r")):
        sim.write_bunches(tmp_path / ("bunches_%s.txt" % tag), header=True)
        sim.write_bunches(tmp_path / ("bunches_%s.txt" % tag))
        sim.write_slices(tmp_path / ("slices_%s.txt" % tag), 2.0)
        sim.dump_particles(tmp_path / ("particles_%s.txt" % tag))
        sim.diagnostics(tmp_path / ("diag_%s.txt" % tag), header=True)
    for name in ("bunches", "slices", "particles"):
        a, b = (tmp_path / ("%s_g.txt" % name)).read_text(), (tmp_path / ("%s_r.txt" % name)).read_text()
        assert len(b.splitlines()) > 2
        assert a == b, name
    dg, dr = [(tmp_path / ("diag_%s.txt" % t)).read_text().splitlines() for t in "gr"]
    assert dg[0] == dr[0]
    vg, vr = np.array(dg[1].split(), dtype=float), np.array(dr[1].split(), dtype=float)
    np.testing.assert_allclose(vg, vr, rtol=1e-4, atol=0)
    assert gpu.field_energy() == pytest.approx(ref.field_energy(), rel=1e-12)


def test_frame_and_axis(tmp_path):
    gpu, ref = g3d_amd.Sim(plasma_deck()), refpy.RefSim(plasma_deck())
    gpu.init(); ref.init()
    gpu.step(); ref.step()
    same_state(gpu, ref, ("F", "rho_lines"))
    for sim, tag in ((gpu, "g"), (ref, "r")):
        sim.write_frame(tmp_path / ("fields_%s.bin" % tag))
        sim.write_frame(tmp_path / ("frame_%s.bin" % tag), comps=0b1000101)
        sim.write_axis(tmp_path / ("axis_%s.txt" % tag))
    for name in ("fields_%s.bin", "frame_%s.bin", "axis_%s.txt"):
        a, b = (tmp_path / (name % "g")).read_bytes(), (tmp_path / (name % "r")).read_bytes()
        assert len(b) > 100
        assert a == b, name


def test_checkpoint_both_ways(tmp_path):
    gpu, ref = g3d_amd.Sim(plasma_deck()), refpy.RefSim(plasma_deck())
    gpu.init(); ref.init()
    for _ in range(2):
        gpu.step(); ref.step()
    step = gpu.info()["istep"]
    # engine -> reference
    (tmp_path / "g").mkdir()
    gpu.write_checkpoint(tmp_path / "g")
    r2 = refpy.RefSim(plasma_deck())
    r2.read_checkpoint(tmp_path / "g", step)
    for n in CKPT + ["js0x", "js0y", "js0z", "F"]:
        np.testing.assert_array_equal(r2.array(n), gpu.array(n), err_msg=n)
    np.testing.assert_array_equal(r2.array("parts"), gpu.array("parts"))
    for k in ("istep", "s", "X0", "Y0", "beta_b", "gm2_b"):
        assert r2.info()[k] == gpu.info()[k], k
    # reference -> engine
    (tmp_path / "r").mkdir()
    ref.write_checkpoint(tmp_path / "r")
    g2 = g3d_amd.Sim(plasma_deck())
    g2.read_checkpoint(tmp_path / "r", step)
    for n in CKPT + ["js0x", "js0y", "js0z", "F"]:
        np.testing.assert_array_equal(g2.array(n), ref.array(n), err_msg=n)


def test_restart_continues_the_run(tmp_path):
    """init + 4 steps, against init + 2 steps, a checkpoint, a fresh simulation reading it and 2 more steps"""
    a = g3d_amd.Sim(bunch_deck())
    a.init()
    for _ in range(2):
        a.step()
    a.write_checkpoint(tmp_path)
    step = a.info()["istep"]
    for _ in range(2):
        a.step()
    b = g3d_amd.Sim(bunch_deck())
    b.read_checkpoint(tmp_path, step)
    for _ in range(2):
        b.step()
    ns, nk = a.info()["ns"], a.info()["nk"]
    assert group_rel_l2([b.array(n) for n in SPEC], [a.array(n) for n in SPEC], ns) <= 1e-12
    assert F_rel_l2(b.array("F"), a.array("F"), ns, nk) <= 1.2e-7
    np.testing.assert_array_equal(b.array("parts")["alive"], a.array("parts")["alive"])
    pa, pb = a.array("parts"), b.array("parts")
    assert np.abs(pa["pz"] - pb["pz"]).max() <= 1e-12 * np.abs(pa["pz"]).max()


def test_checkpoint_refuses_other_grid(tmp_path):
    a = g3d_amd.Sim(bunch_deck())
    a.init()
    a.write_checkpoint(tmp_path)
    b = g3d_amd.Sim(bunch_deck(nz=32))
    with pytest.raises(RuntimeError, match="differ from the deck"):
        b.read_checkpoint(tmp_path, 0)
