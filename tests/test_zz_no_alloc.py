"""g3d_stream_spectra_static, the entry point that allocates nothing (its workspace is a file-scope device array),
against g3d_stream_spectra with a caller's workspace on the same input: the same spectra, bit for bit, on two calls
in a row, the second on a smaller grid (the static workspace keeps the first call's contents); and a grid above the
workspace's limit is refused with a clear error."""
import ctypes as C

import numpy as np
import pytest

import g3d_amd
from decks import plasma_deck

pytestmark = pytest.mark.gpu
D = g3d_amd.DevArray


def spectra_pair(n, seed):
    deck = plasma_deck(nx=n, ny=n, nz=8, smooth=2.0)
    sim = g3d_amd.Sim(deck)
    tab = {k: D(sim.array(k)) for k in ("kx", "ky", "kxs", "kys", "filt", "filty", "hcg", "wlo", "whi")}
    plan = g3d_amd.FftPlan(n, n)
    info = sim.info()
    nk = info["nk"]
    spec = g3d_amd.SpecCtx(nx=n, ny=n, ns=info["ns"], N=info["N"], a=0, nk=nk, a2=4.0, keep0=1, plan=plan.h.value,
                           **{k: v.ptr.value for k, v in tab.items()})
    jsr = np.random.default_rng(seed).standard_normal(5 * nk)
    ja, jb = D(jsr), D(jsr)
    sa, sb = D(nbytes=3 * nk * 16), D(nbytes=3 * nk * 16)
    work = D(nbytes=6 * nk * 16)
    g3d_amd.kernel_call("g3d_stream_spectra", C.byref(spec), ja.ptr, sa.ptr, work.ptr)
    g3d_amd.kernel_call("g3d_stream_spectra_static", C.byref(spec), jb.ptr, sb.ptr)
    a, b = sa.get(np.complex128), sb.get(np.complex128)
    assert np.abs(a).max() > 0
    np.testing.assert_array_equal(b, a)
    np.testing.assert_array_equal(jb.get(np.float64), np.zeros(5 * nk))    # the current is consumed, as with a workspace
    return spec, plan, tab


def test_static_workspace_equals_caller_workspace():
    spectra_pair(128, 1)
    spectra_pair(32, 2)           # smaller: the workspace still holds the first call's larger spectra


def test_static_workspace_limit():
    spec, plan, tab = spectra_pair(16, 3)
    big = g3d_amd.SpecCtx(nx=512, ny=512, ns=1, N=2, a=0, nk=512 * 512, plan=plan.h.value)
    rc = g3d_amd.lib().g3d_stream_spectra_static(C.byref(big), None, None)
    assert rc != 0
    assert b"above" in g3d_amd.lib().g3d_last_error() or b"outside" in g3d_amd.lib().g3d_last_error()
