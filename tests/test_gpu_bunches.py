"""Several bunches (the reference's bunch, bunch2, bunch3 blocks) against the reference's own CPU build (oracle/_ref,
one thread), on the Fourier and the free-hand y grid: each bunch drawn from its own seed with its species code
(0, -1, -2) equals the reference's particle for particle; the start field, the co-moving field of each bunch with
its own beta and the forward fields summed (src/sim_bunches.cpp:218-248), is within the FP64 bars, and so are the
particles and fields after two steps; bunches.txt has a line per bunch, the same text as the reference's."""
import numpy as np
import pytest

import g3d_amd
import refpy
from decks import bunch_deck, particle_parity
from test_gpu_engine import compare_fields
from test_gpu_freehand import fh_bunch_deck

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not refpy.available(), reason="oracle/_ref not built")]


def three(freehand):
    d = fh_bunch_deck(nmacro=2048) if freehand else bunch_deck(nmacro=2048)
    Lw = float(d.zeta[-1])
    d.add_bunch("witness", 2e-12, 200.0, 1.0, 1.5, 1.0, 0.3 * Lw, 1024, 2, emitx=1e-3, espread=1e-3)
    d.add_bunch("", 1e-12, 50.0, 1.5, 1.5, 2.0, 0.15 * Lw, 512, 9)
    return d


@pytest.mark.parametrize("freehand", [False, True])
def test_three_bunches_against_reference(freehand, tmp_path):
    refpy.set_threads(1)
    gpu, ref = g3d_amd.Sim(three(freehand)), refpy.RefSim(three(freehand))
    gpu.init(); ref.init()
    a, b = gpu.array("parts"), ref.array("parts")
    assert len(b) > 3000
    np.testing.assert_array_equal(np.unique(b["sp"]), [-2, -1, 0])
    for f in ("x", "y", "zeta", "px", "py", "pz", "w", "sp"):
        np.testing.assert_array_equal(a[f], b[f], err_msg=f)
    for k in ("beta_b", "gm2_b"):
        assert gpu.info()[k] == ref.info()[k], k
    compare_fields(gpu, ref, groups=("fields",))
    for _ in range(2):
        gpu.step(); ref.step()
    a, b = gpu.array("parts"), ref.array("parts")
    np.testing.assert_array_equal(a["alive"], b["alive"])
    live = b["alive"] == 1
    particle_parity(a[live], b[live], three(freehand))
    compare_fields(gpu, ref, groups=("fields",))
    gpu.write_bunches(tmp_path / "g.txt", header=True)
    ref.write_bunches(tmp_path / "r.txt", header=True)
    tg, tr = (tmp_path / "g.txt").read_text(), (tmp_path / "r.txt").read_text()
    assert "witness" in tr and "bunch3" in tr
    assert tg == tr
