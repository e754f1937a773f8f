"""The deck reader (g3d_deck) against the reference's own (Deck::read, overrides and Config::from_deck of its CPU
build, oracle/_ref): on the same deck text every field of the run configuration, the zeta and y tables and the
output settings are the same numbers, bit for bit; decks the reference refuses are refused; decks that need what
this engine does not have are refused with the key named."""
import ctypes as C
import os

import numpy as np
import pytest

import g3d_deck
import refpy

pytestmark = pytest.mark.skipif(not refpy.available(), reason="oracle/_ref not built")
EXAMPLES = "/root/reference/examples"

FOURIER = """
# Fourier transverse grid, the analytic plane wave and a bunch, every key given
grid {
  lambda0 = 10 um; nx = 32; ny = 16
  Lx = 190.986 um ; Ly = 120 um
  nzeta = 65 ; window_length = 238.732 um
  dz = 15.9155 um; z_end = 2.38732mm ; dz_push = 5 um
  sponge = 10 um; sponge_y = 7 um; sponge_rate = 0.2
  move_x = off ; move_y = yes ; window_lag = 3
}
laser {
  type = plane_wave ; a0 = 0.5 ; wavelength = 9.5 um ; angle = 5 deg ; tau_fwhm = 0.3 ps
  cep = 30 mrad ; polarization = p ; start_fwhm = 3 ; stop_fwhm = 0
}
bunch {
  charge = 1.41764 pC ; energy = 25 MeV ; sigma_x = 7.95775 um ; sigma_y = 6 um ; sigma_z = 7.95775 um
  emittance_x = 0.01 um ; emittance_y = 0.02 um ; energy_spread = 0.001
  zeta_center = 190.986 um ; macroparticles = 2e4 ; seed = 3 ; species = electron ; shot_noise = random
  distribution_r = gaussian ; distribution_z = gaussian ; name = "bunch" ; x0 = 0 um ; y0 = 0 um
}
numerics { self_fields = true ; zeta_filter = false ; k_filter = 0 ; push_phase = 0.02 ; iterations = 3
  vz_min = 0.7 ; vz_beam = 0.95 ; trap_advance = 2 ; predictor = previous ; freeze_bunch = on ; start_relax = no
  box_mean = remove }
output { dir = "run one" ; every = 5 ; beam_every = 7 ; fields = ez, bz ne ; beams = bunch trapped ; axis = true
  particles = false ; frames = true ; slice_length = 2 um ; checkpoint_every = 4 ; walltime_hours = 1.5
  format = native ; xi_stride = 1 ; particle_stride = 1 }
"""

FREEHAND = """
grid {
  lambda0 = 0.8 um
  nx = 16
  Lx = 80 um
  y_grid = table
  y_points = 0, 10, 30 um
  dy_points = 1, 1, 4 um
  y_max = 40 um
  zeta_grid = table
  zeta_points = 0 20 40 um
  dzeta_points = 1 1 0.25 um
  window_length = 60 um
  dz = 4 um
  z_end = 0.1 mm
  move_x = false
}
laser { type = gaussian; a0 = 2; waist = 12 um; tau_fwhm = 25 fs; zeta_center = 48 um; z_focus = 0.5 mm
  polarization = circular ; helicity = -1 ; cep = 0.3 ; pol_angle = 10 deg ; wavelength = 0.81 um }
laser { type = gaussian ; a0 = 0.1 ; waist = 20 um ; tau_fwhm = 40 fs ; zeta_center = 30 um }
laser { type = gaussian ; enabled = false }
plasma {
  density = 3.5e18 cm^-3 ; z_points = 0, 100, 200 um ; z_values = 0, 1, 0.5
  profile = parabolic ; channel_depth = 0.3 ; r_channel = 15 um ; r_max = 30 um
  ppc_x = 3 ; ppc_y = 1 ; smooth_length = 0.25 um ; free_electrons = true
}
ion { element = n ; charge_state = 5 ; density = 3.5e16 cm^-3 ; ionization = adk ; mobile = false }
ion { element = He ; charge_state = 2 ; fraction = 0.5 }
ion { mass = 2 u ; ionization_potentials = 10, 20 eV ; charge_state = 1 ; ionization = adk ; name = Xx
  max_charge_state = 2 ; fraction = 0.1 }
ion { element = H ; fraction = 0.05 ; mobile = true }
"""

BUNCHES = """
grid { nx = 16 ; ny = 16 ; Lx = 50 um ; nzeta = 33 ; window_length = 30 um ; dz = 1 um ; z_end = 10 um }
bunch { charge = 10 pC ; gamma = 1000 ; sigma_x = 2 um ; sigma_z = 2 um ; zeta_center = 20 um
  macroparticles = 1000 }
bunch { charge = 3 pC ; energy = 50 MeV ; sigma_x = 1 um ; sigma_y = 1.5 um ; sigma_z = 1 um ; zeta_center = 8 um
  macroparticles = 500 ; name = witness ; emittance_x = 0.1 um ; energy_spread = 0.01 }
bunch { charge = 1 pC ; gamma = 200 ; sigma_x = 1 um ; sigma_z = 1 um ; seed = 9 }
"""

MINIMAL = """
grid { nx = 8 ; ny = 8 ; Lx = 50 um ; nzeta = 9 ; window_length = 30 um ; dz = 1 um ; z_end = 10 um }
plasma { density = 0.01 n_c }
"""


def plain(v):
    """a ctypes array as a list (compared by value)"""
    return list(v) if isinstance(v, C.Array) else v


def compare(path, overrides=()):
    mine = g3d_deck.read_deck(path, overrides)
    cfg, zeta, ynode, out = refpy.deck_config(path, overrides)
    c = mine.c
    for name, typ in type(cfg)._fields_:
        if name in ("zeta", "ynode"):
            continue
        a, b = getattr(c, name), getattr(cfg, name)
        if name == "glasers":
            for k in range(cfg.n_glasers):
                for f, _ in type(a[k])._fields_:
                    assert plain(getattr(a[k], f)) == plain(getattr(b[k], f)), "glasers[%d].%s" % (k, f)
        elif name == "ions":
            for k in range(cfg.n_ions):
                for f, _ in type(a[k])._fields_:
                    va, vb = getattr(a[k], f), getattr(b[k], f)
                    if f == "ip_eV":
                        va, vb = list(va)[:a[k].n_ip], list(vb)[:b[k].n_ip]
                    assert va == vb, "ions[%d].%s: %r %r" % (k, f, va, vb)
        elif name == "bunch_more":
            for k in range(cfg.n_bunch_more):
                for f, _ in type(a[k])._fields_:
                    assert plain(getattr(a[k], f)) == plain(getattr(b[k], f)), "bunch_more[%d].%s" % (k, f)
        elif name in ("pz_points", "pz_values"):
            assert list(a)[:cfg.n_pz] == list(b)[:cfg.n_pz], name
        else:
            assert plain(a) == plain(b), "%s: %r (reader) against %r (reference)" % (name, a, b)
    np.testing.assert_array_equal(mine.zeta, zeta)
    np.testing.assert_array_equal(mine.ynode, ynode)
    for k, v in out.items():
        w = mine.output[k]
        assert (w == v) if k == "dir" else (float(w) == v), "output.%s: %r %r" % (k, w, v)
    return mine


def write(tmp_path, text, name="d.deck"):
    p = tmp_path / name
    p.write_text(text)
    return p


@pytest.mark.parametrize("text", [FOURIER, FREEHAND, MINIMAL, BUNCHES], ids=["fourier", "freehand", "minimal",
                                                                        "bunches"])
def test_reader_matches_reference(tmp_path, text):
    compare(write(tmp_path, text))


def test_overrides(tmp_path):
    p = write(tmp_path, FOURIER)
    d = compare(p, ["grid.nx=64", "laser.a0 = 0.25", "bunch.seed=11", "numerics.iterations=1", "output.every=2"])
    assert d.c.nx == 64 and d.c.seed == 11
    compare(write(tmp_path, MINIMAL, "m.deck"), ["plasma.density=3e18 cm^-3", "bunch.charge=1 pC",
                                                  "bunch.gamma=100", "bunch.sigma_x=1 um", "bunch.sigma_z=2 um"])


@pytest.mark.skipif(not os.path.isdir(EXAMPLES), reason="the reference's example decks are absent")
@pytest.mark.parametrize("name", ["beam_wake_frozen", "ion_collapse", "ionization_injection", "lwfa_plasma"])
def test_reference_examples(name):
    compare(os.path.join(EXAMPLES, name + ".deck"))


@pytest.mark.skipif(not os.path.isdir(EXAMPLES), reason="the reference's example decks are absent")
def test_example_with_fawley_noise_refused():
    with pytest.raises(g3d_deck.DeckError, match="noise_wavelength|shot_noise|beamlet"):
        g3d_deck.read_deck(os.path.join(EXAMPLES, "oblique_plane_wave.deck"))


BAD = {
    "unknown key": MINIMAL + "numerics { colour = red }",
    "block twice": MINIMAL + "grid { nx = 8 }",
    "unclosed": MINIMAL + "numerics { iterations = 2",
    "no unit": MINIMAL.replace("Lx = 50 um", "Lx = 50"),
    "wrong unit kind": MINIMAL.replace("Lx = 50 um", "Lx = 50 fs"),
    "unknown unit": MINIMAL.replace("Lx = 50 um", "Lx = 50 furlong"),
    "unit on a count": MINIMAL.replace("nx = 8", "nx = 8 um"),
    "non-integer": MINIMAL.replace("nx = 8", "nx = 8.5"),
    "nx too small": MINIMAL.replace("nx = 8", "nx = 2"),
    "bad flag": MINIMAL + "numerics { k_filter = maybe }",
    "vz order": MINIMAL + "numerics { vz_min = 0.95 }",
    "ion without plasma": MINIMAL.replace("plasma { density = 0.01 n_c }", "") + "ion { element = H }",
    "ion charge": MINIMAL + "ion { element = He ; charge_state = 3 }",
    "ion fraction and density": MINIMAL + "ion { element = H ; fraction = 1 ; density = 1e18 cm^-3 }",
    "neutral without density": MINIMAL + "ion { element = He ; charge_state = 0 ; ionization = adk }",
    "too much ion charge": MINIMAL + "ion { element = He ; charge_state = 2 ; fraction = 1.5 }",
    "gamma and energy": MINIMAL + "bunch { charge = 1 pC ; gamma = 10 ; energy = 5 MeV ; sigma_x = 1 um ; "
                                  "sigma_z = 1 um }",
    "table y with ny": MINIMAL.replace("ny = 8", "ny = 8 ; y_grid = table"),
    "box_mean remove, free-hand": FREEHAND + "numerics { box_mean = remove }",
    "key of a disabled laser": MINIMAL + "laser { type = gaussian ; enabled = false ; a0 = 9 }",
    "list with two units": MINIMAL.replace("density = 0.01 n_c", "density = 0.01 n_c ; z_points = 0, 1 um, 2 mm ; "
                                           "z_values = 0, 1, 1"),
}


@pytest.mark.parametrize("case", sorted(BAD))
def test_both_refuse(tmp_path, case):
    p = write(tmp_path, BAD[case])
    with pytest.raises(RuntimeError):
        refpy.deck_config(p)
    with pytest.raises(g3d_deck.DeckError):
        g3d_deck.read_deck(p)


UNSUPPORTED = {
    "expression": (MINIMAL.replace("Lx = 50 um", "Lx = 2 * 25 um"), "expressions"),
    "constants": (MINIMAL + "constants { w = 5 um }", "constants"),
    "fifth bunch": (MINIMAL + "bunch { charge = 1 pC ; gamma = 10 ; sigma_x = 1 um ; sigma_z = 1 um }\n" * 5,
                    "at most 4"),
    "flat-top": (MINIMAL + "bunch { charge = 1 pC ; gamma = 10 ; sigma_x = 1 um ; distribution_z = flattop ; "
                 "length = 3 um }", "distribution_z"),
    "envelope": (MINIMAL + "laser { type = gaussian ; a0 = 1 ; waist = 5 um ; envelope_function = 1 }",
                 "envelope_function"),
    "file laser": (MINIMAL + "laser { type = file ; file = x.h5 }", "file"),
    "undulator": (MINIMAL + "undulator { period = 1 cm ; K = 1 ; periods = 10 }", "undulator"),
    "openpmd": (MINIMAL + "output { format = openpmd }", "native"),
    "heavy element": (MINIMAL + "ion { element = Ar ; charge_state = 1 }", "16 ionization levels"),
}


@pytest.mark.parametrize("case", sorted(UNSUPPORTED))
def test_unsupported_named(tmp_path, case):
    text, what = UNSUPPORTED[case]
    with pytest.raises(g3d_deck.DeckError, match=what):
        g3d_deck.read_deck(write(tmp_path, text))


def test_parse_with_unit_forms():
    P = g3d_deck.parse_with_unit
    assert P("3um", "length") == 3e-6 and P(" 2.5e-3 mrad", "angle") == 2.5e-6
    assert P("0.1", "angle") == 0.1 and P("-4", None) == -4.0 and P(".5 keV", "energy") == 500.0
    assert P("1 T/m", "gradient") == 1.0 and P("2 cm^-3", "density") == 2e6
    for bad, kind in (("um", "length"), ("3", "length"), ("3 um", None), ("3 fs", "length")):
        with pytest.raises(g3d_deck.DeckError):
            P(bad, kind)
