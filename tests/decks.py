"""Small decks of the parity tests (code units, lambda0 = 2 pi um so that 1 um = 1/k0) and the comparison metrics."""
import numpy as np

import g3d_amd

LAMBDA0 = 2 * np.pi * 1e-6


def plasma_deck(nx=32, ny=32, nz=48, ppc=2, sponge=6.0, passes=2, smooth=0.0, a0=1.0, n0=1e-2, extrap=1):
    """A reduced LWFA deck: uniform plasma, Gaussian grid laser (config 1 of the benchmark, shrunk)."""
    h = 0.5
    d = g3d_amd.Deck(LAMBDA0).grid(nx, ny, nz, 60.0, 60.0, (nz - 1) * h)
    d.set(dz=5.0, z_end=1000.0, plasma_on=1, n0=n0, ppc_x=ppc, ppc_y=ppc, move_y=0, sponge=sponge, sponge_y=sponge,
          iterations=passes, smooth_length=smooth, predictor_extrapolate=extrap)
    d.add_grid_laser(a0=a0, waist=8.0, tau=3.0, zeta_c=(nz - 1) * h * 0.6)
    return d


def bunch_deck(nx=32, ny=32, nz=64, nmacro=4096, laser=True, sponge=4.0):
    """A reduced config-0 deck: electron bunch at gamma 1000 plus the oblique plane-wave laser, vacuum."""
    h = 0.5
    d = g3d_amd.Deck(LAMBDA0).grid(nx, ny, nz, 40.0, 40.0, (nz - 1) * h)
    Lw = (nz - 1) * h
    d.set(dz=20.0, z_end=1000.0, sponge=sponge, sponge_y=sponge, move_x=1, move_y=1,
          charge=11.4014089e-12 * nmacro / 262144, gamma=1000.0, sigx=2.0, sigy=2.0, sigz=4.0, emitx=2e-4, emity=2e-4,
          espread=1e-3, zeta_center=0.5 * Lw, nmacro=nmacro, seed=1)
    if laser:
        d.set(laser_on=1, laser_pol=1, a0=1.0, theta=2e-3, tau=30.0, start_fwhm=0.0, stop_fwhm=0.0)
    return d


def slice_rel_l2(a, b, nplanes, floor=1e-6):
    """max over planes of ||a - b|| / ||b||; planes whose norm is below floor * (largest plane norm) use that floor."""
    a = np.asarray(a).reshape(nplanes, -1)
    b = np.asarray(b).reshape(nplanes, -1)
    nb = np.linalg.norm(b, axis=1)
    den = np.maximum(nb, floor * max(nb.max(), 1e-300))
    return float((np.linalg.norm(a - b, axis=1) / den).max())


def group_rel_l2(A, B, nplanes, floor=1e-6):
    """Relative L2 per plane over a group of field components taken together (the field vector of a slice):
    max_n ||A_n - B_n|| / ||B_n||, with the planes below floor * (largest plane norm) measured against that floor.
    A, B: lists of arrays of nplanes planes each."""
    a = np.concatenate([np.asarray(x).reshape(nplanes, -1) for x in A], axis=1)
    b = np.concatenate([np.asarray(x).reshape(nplanes, -1) for x in B], axis=1)
    return slice_rel_l2(a, b, nplanes, floor)


def F_rel_l2(Fa, Fb, ns, nk):
    """relative L2 per slice of the real fields F [6][ns+2][nk] (E and B of a slice together), own slices"""
    A = np.asarray(Fa, dtype=np.float64).reshape(6, ns + 2, nk)[:, :ns]
    B = np.asarray(Fb, dtype=np.float64).reshape(6, ns + 2, nk)[:, :ns]
    return group_rel_l2([A[c] for c in range(6)], [B[c] for c in range(6)], ns)


def particle_parity(a, b, deck, tol=1e-12):
    """particle lists a (GPU) and b (reference) of the same length and order: cells bit-exact (x cell, y cell of
    the Fourier nodes or of the free-hand node table, zeta cell), positions and the momentum vectors within tol
    relative L2 over the list (the momentum bar of the FP64 form)"""
    dx = deck.c.Lx / deck.c.nx
    np.testing.assert_array_equal(np.floor(a["x"] / dx), np.floor(b["x"] / dx))
    if deck.c.y_fourier:
        dy = deck.c.Ly / deck.c.ny
        np.testing.assert_array_equal(np.floor(a["y"] / dy), np.floor(b["y"] / dy))
    else:
        np.testing.assert_array_equal(np.searchsorted(deck.ynode, a["y"]), np.searchsorted(deck.ynode, b["y"]))
    np.testing.assert_array_equal(np.searchsorted(deck.zeta, a["zeta"]), np.searchsorted(deck.zeta, b["zeta"]))
    for f in ("x", "y", "zeta"):
        assert np.linalg.norm(a[f] - b[f]) <= tol * np.linalg.norm(b[f]), f
    p = np.sqrt(b["px"] ** 2 + b["py"] ** 2 + b["pz"] ** 2)
    dp = np.sqrt((a["px"] - b["px"]) ** 2 + (a["py"] - b["py"]) ** 2 + (a["pz"] - b["pz"]) ** 2)
    assert np.linalg.norm(dp) <= tol * np.linalg.norm(p), np.linalg.norm(dp) / np.linalg.norm(p)


def config2_deck(mixed=0, nz=1024):
    """The laser-wakefield deck at full size (256 x 256 modes, 1024 slices of 0.25/k0, 262144 streaming electrons
    per slice, two passes, Gaussian grid laser a0 = 2, w0 = 30/k0, c tau = 15/k0 at zeta = 200/k0) with the trailing
    witness bunch (2^20 electrons, sigma_r = 1.5/k0, sigma_zeta = 3/k0, 0.005 n_c peak = 2.40 pC, gamma = 200, at
    zeta = 100/k0 in the first bucket behind the laser: the plasma wavelength is 2 pi / sqrt(1e-3) = 199/k0)."""
    d = g3d_amd.Deck(LAMBDA0)
    d.c.nx = 256; d.c.Lx = 240.0
    d.y_uniform(256, 240.0)
    d.zeta_uniform(nz, (nz - 1) * 0.25)
    d.set(dz=20.0, z_end=4000.0, sponge=24.0, sponge_y=24.0, move_x=0, move_y=0, plasma_on=1, n0=1e-3,
          ppc_x=2, ppc_y=2, smooth_length=0.0, iterations=2, mixed=mixed)
    d.add_grid_laser(a0=2.0, waist=30.0, tau=15.0, zeta_c=200.0, z_focus=0.0)
    d.set(charge=2.40e-12, gamma=200.0, sigx=1.5, sigy=1.5, sigz=3.0, emitx=1e-3, emity=1e-3, espread=0.0,
          zeta_center=100.0, nmacro=1 << 20, seed=7)
    return d


def bunch_moments(p):
    """The beam moments of the reference's bunches.txt for the live particles of a list, one group
    (src/sim_bunch_diag.cpp:9-61): the weighted first moments, then the second moments centred on them (two passes);
    returns the total weight, <gamma> and the normalised emittances eps_nx, eps_ny (code units)."""
    live = p[p["alive"] == 1]
    w = live["w"]
    W = w.sum()
    g = np.sqrt(1 + live["px"] ** 2 + live["py"] ** 2 + live["pz"] ** 2)
    x, y, ux, uy = live["x"], live["y"], live["px"], live["py"]
    mx, my, mux, muy = [(w * v).sum() / W for v in (x, y, ux, uy)]
    dx, dy, dux, duy = x - mx, y - my, ux - mux, uy - muy
    var = lambda a, b: (w * a * b).sum() / W
    ex = np.sqrt(max(0.0, var(dx, dx) * var(dux, dux) - var(dx, dux) ** 2))
    ey = np.sqrt(max(0.0, var(dy, dy) * var(duy, duy) - var(dy, duy) ** 2))
    return W, float((w * g).sum() / W), float(ex), float(ey)


def F_slice_errors(Fa, Fb, ns, nk):
    """relative L2 of F per slice (E and B of the slice together) over every own slice, and the slice norms"""
    A = np.asarray(Fa, dtype=np.float64).reshape(6, ns + 2, nk)[:, :ns]
    B = np.asarray(Fb, dtype=np.float64).reshape(6, ns + 2, nk)[:, :ns]
    nb = np.sqrt((B ** 2).sum(axis=(0, 2)))
    de = np.sqrt(((A - B) ** 2).sum(axis=(0, 2)))
    return de / np.maximum(nb, 1e-300), nb


def config0_deck():
    """Config 0 at full size: the vacuum deck of 64 x 64 modes over 80/k0, 256 slices of 0.5/k0, 50 steps of 20/k0;
    the bunch of 2^18 electrons (11.4014 pC: peak 0.01 n_c for sigma = 2, 2, 4/k0), gamma 1000, energy spread 1e-3,
    emittance 2e-4 um (u_x rms 1e-4), seed 1; the plane wave a0 = 1, p-polarised, tilted 2 mrad, c tau = 30/k0"""
    d = g3d_amd.Deck(LAMBDA0)
    d.c.nx = 64; d.c.Lx = 80.0
    d.y_uniform(64, 80.0)
    d.zeta_uniform(256, 255 * 0.5)
    d.set(dz=20.0, z_end=1000.0, sponge=8.0, sponge_y=8.0, move_x=1, move_y=1, charge=11.4014089e-12, gamma=1000.0,
          sigx=2.0, sigy=2.0, sigz=4.0, emitx=2e-4, emity=2e-4, espread=1e-3, zeta_center=0.5 * 255 * 0.5,
          nmacro=1 << 18, seed=1, laser_on=1, laser_pol=1, a0=1.0, theta=2e-3, tau=30.0, start_fwhm=0.0,
          stop_fwhm=0.0)
    return d


def config4_params(seed=4):
    """the parameters of config 4: the 32 laser amplitudes a0(zeta) = 1.5 + N(0, 0.05) (seeded) and the 16 values of
    the witness current profile (ones)"""
    rng = np.random.default_rng(seed)
    return 1.5 + 0.05 * rng.standard_normal(32), np.ones(16)


def config4_deck(env=None, prof=None, n=128, nz=512, nwit=1 << 18, steps=20, mixed=0):
    """Config 4: 128 x 128 modes over 120/k0, 512 slices of 0.5/k0, 20 steps of 10/k0; uniform plasma at 2e-3 n_c,
    2 x 1 electrons per cell; the Gaussian driver (waist 25/k0) whose amplitude along zeta is the piecewise-linear
    a0(zeta) through 32 values 2/k0 apart (its front 10/k0 behind the window's front); the witness of nwit electrons
    at gamma 100 (sigma_r 2/k0, 5 pC) whose current profile is the piecewise-linear profile through 16 values over
    zeta 90 ... 120/k0.  Smaller n, nz and nwit give the same deck on a coarser grid (the tests against the
    reference)."""
    e0, p0 = config4_params()
    env = e0 if env is None else env
    prof = p0 if prof is None else prof
    h = 0.5 * 512 / nz
    Lw = (nz - 1) * h
    d = g3d_amd.Deck(LAMBDA0)
    d.c.nx = n; d.c.Lx = 120.0
    d.y_uniform(n, 120.0)
    d.zeta_uniform(nz, Lw)
    d.set(dz=10.0, z_end=10.0 * steps, sponge=8.0, sponge_y=8.0, move_x=0, move_y=0, plasma_on=1, n0=2e-3,
          ppc_x=2, ppc_y=1, iterations=2, mixed=mixed, charge=5e-12, gamma=100.0, sigx=2.0, sigy=2.0, sigz=5.0,
          emitx=0.0, emity=0.0, espread=0.0, nmacro=nwit, seed=11)
    z_front = Lw - 10.0
    d.add_grid_laser(a0=1.0, waist=25.0, tau=1.0, zeta_c=z_front - 31.0, z_focus=0.0)
    d.set_envelope(0, list(env), z_front - 62.0, 2.0)
    d.set_profile(list(prof), 90.0 * Lw / 255.5, 120.0 * Lw / 255.5)
    return d
