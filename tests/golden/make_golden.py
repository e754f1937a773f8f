"""Writes the golden vectors of the per-slice march from the reference's own CPU build (oracle/_ref, one thread):
small seeded inputs and the reference's outputs, one .npz per function (inputs and expected outputs only).

Run in the build container (the reference tree must be present to build oracle/_ref):
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "gem-pic-3d_amd"))
import refpy  # noqa: E402
import golden_inputs as gi  # noqa: E402


def main(only=None):
    refpy.set_threads(1)
    out = {}
    # batched FFT (src/fftmany.hpp): 2D planes and rows, both directions
    x = gi.fft_input()
    out["fft"] = dict(x=x, fwd2d=refpy.fft_many(8, 16, True, x, 0), inv2d=refpy.fft_many(8, 16, True, x, 1),
                      fwd1d=refpy.fft_many(8, 16, False, x, 0), inv1d=refpy.fft_many(8, 16, False, x, 1))
    # march of random spectra (src/sim_fields.cpp:7-39)
    deck = gi.small_deck()
    ref = refpy.RefSim(deck)
    arr = gi.march_input(ref)
    for k, v in arr.items():
        ref.set(k, v)
    ref.call("march", gi.MARCH_D)
    out["march"] = dict(**{"in_" + k: v for k, v in arr.items()},
                        **{"out_" + k: ref.array(k) for k in ("Rx", "Ry", "Ez", "Bz", "Lx", "Ly")})
    # march of the k_x lines of a free-hand y grid (src/sim_kernels.hpp:213-313)
    deck = gi.small_fh_deck()
    ref = refpy.RefSim(deck)
    arr = gi.march_input(ref)
    for k, v in arr.items():
        ref.set(k, v)
    ref.call("march", gi.MARCH_D)
    out["march_fh"] = dict(ynode=deck.ynode, **{"in_" + k: v for k, v in arr.items()},
                           **{"out_" + k: ref.array(k) for k in ("Rx", "Ry", "Ez", "Bz", "Lx", "Ly")})
    # streaming current spectra (src/sim_plasma.cpp:94-117)
    deck = gi.small_deck(smooth=1.5)
    ref = refpy.RefSim(deck)
    J = gi.jsr_input()
    ref.set("jsr", J)
    ref.call("stream_spectra")
    out["stream_spectra"] = dict(J=J, jsn=ref.array("jsn"), a2=1.5 ** 2)
    # one streaming half-cell pass (src/sim_stream.hpp:104-171) in random fields
    deck = gi.small_deck()
    ref = refpy.RefSim(deck)
    ref.call("init")
    F, P = gi.stream_input(ref)
    ref.set("F", F)
    ref.set("strm", P)
    ref.call("stream_pass", gi.STREAM_N, 0, 0.0, 5.0)
    out["stream_pass"] = dict(F=F, parts_in=P, parts_out=ref.array("strm"), jsr=ref.array("jsr"),
                              X0=ref.info()["X0"], Y0=ref.info()["Y0"])
    # RK4 push of bunch particles in random fields and the oblique plane wave (src/sim_kernels.hpp:98-149)
    deck = gi.bunch_small_deck()
    ref = refpy.RefSim(deck)
    ref.call("init")
    F, P = gi.push_input(ref)
    ref.set("F", F)
    ref.set_particles(P)
    ref.call("push", 0.0, 20.0)
    out["push"] = dict(F=F, parts_in=P, parts_out=ref.array("parts"), X0=ref.info()["X0"], Y0=ref.info()["Y0"])
    # ADK levels of nitrogen (src/ionization.cpp:30-53), code units of lambda0 = 0.8 um
    out["adk"] = dict(ip=gi.N_IP, table=gi.ref_adk_table())
    # SplitMix64 finaliser and the draw keys (src/ionization.hpp:37-46)
    keys = np.array([0, 1, 1234567, 0x9e3779b97f4a7c15, 2 ** 63 + 12345, 2 ** 64 - 1], dtype=np.uint64)
    out["mix64"] = dict(keys=keys, mix=np.array([refpy.lib().ref_mix64(int(k)) for k in keys], dtype=np.uint64),
                        u01=np.array([refpy.lib().ref_uniform01(int(k)) for k in keys]))
    # one streaming pass with ionizing nitrogen, mobile and fixed, from the neutral atom (src/sim_stream.hpp:67-171):
    # the charge states, the freed electrons (in list order on one thread) and the deposit
    deck = gi.ion_deck()
    ref = refpy.RefSim(deck)
    ref.call("init")
    F, P, cap = gi.ion_stream_input(ref)
    seed, step, n, pas = gi.ION_KEY
    ref.set("F", F)
    ref.set("strm", P)
    ref.call("set_s", 0.0, step)
    ref.call("reserve_born", cap)
    ref.call("stream_pass", n, pas, 0.0, 5.0)
    L = refpy.lib()
    key = L.ref_key_slice(L.ref_key_step(seed, step), n, pas)
    out["stream_ions"] = dict(F=F, parts_in=P, parts_out=ref.array("strm"), jsr=ref.array("jsr"), born=ref.array("born"),
                              cap=cap, key=np.uint64(key), X0=ref.info()["X0"], Y0=ref.info()["Y0"])
    # one streaming pass on a free-hand y grid and a free-hand zeta table, with the plane wave (src/sim_stream.hpp:
    # 33-171, the table locates of src/grid.hpp:35-46): walls, nodes and half-nodes of the deposit
    deck = gi.fh_stream_deck()
    ref = refpy.RefSim(deck)
    ref.call("init")
    F, P = gi.fh_stream_input(ref, deck)
    ref.set("F", F)
    ref.set("strm", P)
    ref.call("stream_pass", gi.FH_STREAM_N, 0, 0.0, 5.0)
    out["stream_fh"] = dict(F=F, parts_in=P, parts_out=ref.array("strm"), jsr=ref.array("jsr"), ynode=deck.ynode,
                            zeta=deck.zeta, X0=ref.info()["X0"], Y0=ref.info()["Y0"])
    # the beam deposit of currents() (src/sim_particles.cpp:56-107): jw on one thread, in list order
    deck = gi.bunch_small_deck()
    ref = refpy.RefSim(deck)
    P = gi.deposit_input()
    ref.set_particles(P)
    ref.call("currents")
    out["bunch_deposit"] = dict(parts=P, jw=ref.array("jw"))
    # the key chain of the ionization draws (src/ionization.hpp:44-46)
    ks, kl, kp = [], [], []
    for seed, step, n, pas in gi.KEY_CASES:
        a = L.ref_key_step(seed, step)
        b = L.ref_key_slice(a, n, pas)
        ks.append(a); kl.append(b); kp.append([L.ref_key_particle(b, int(i)) for i in gi.KEY_IDS])
    out["keys"] = dict(seed=np.array([c[0] for c in gi.KEY_CASES], dtype=np.uint64),
                       step=np.array([c[1] for c in gi.KEY_CASES], dtype=np.int64),
                       n=np.array([c[2] for c in gi.KEY_CASES], dtype=np.int64),
                       pas=np.array([c[3] for c in gi.KEY_CASES], dtype=np.int64), ids=gi.KEY_IDS,
                       key_step=np.array(ks, dtype=np.uint64), key_slice=np.array(kl, dtype=np.uint64),
                       key_particle=np.array(kp, dtype=np.uint64))
    for name, d in out.items():
        if only and name not in only:
            continue
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
        print("wrote", name, os.path.getsize(os.path.join(HERE, name + ".npz")), "bytes")


if __name__ == "__main__":
    main(sys.argv[1:])
