"""Benchmark of the per-slice march on one MI355X: the laser-wakefield deck with a trailing witness bunch
(256 x 256 Fourier modes, 1024 uniform zeta slices, h = 0.25/k0, uniform plasma 1e-3 n_c streamed as 2 x 2
electrons per cell per slice = 262144 per slice, two passes per half-cell, Gaussian grid laser a0 = 2, w0 = 30/k0,
c tau = 15/k0 at zeta = 200/k0, sponge 24/k0, dz = 20/k0; witness of 2^20 electrons, sigma_r = 1.5/k0,
sigma_zeta = 3/k0, 0.005 n_c peak = 2.40 pC, gamma = 200, at zeta = 100/k0 in the first bucket behind the laser).
A step is one s-step of the march: the RK4 push of the witness, its deposit and current spectra, the 1023
half-cells x 2 passes of the streaming plasma, sponge and real fields.  It runs in the FP64 form and in the
FP32-mixed form from the same start.  Metric: live pushes per second (streaming particles through a half-cell in a
pass, counted on the device, plus witness particles advanced by a step) including deposit and gather; the line
reports the mixed form, with its per-slice field error against FP64 and the witness's energy-gain and emittance
errors from the reference's centred moments.

--workload freehand: config 3 instead (512 Fourier x modes by 386 free-hand y cells between conducting walls,
2048 slices on a zeta table from 0.1 to 0.5/k0, electrons 2 x 1 per cell, mobile hydrogen, nitrogen with ADK
ionization, a0 = 2.5 Gaussian laser); metric: streaming-particle pushes per second (ions and freed electrons
included), the plain FP64 form.

--workload vacuum: the vacuum deck instead (64 x 64 modes, 256 slices, h = 0.5/k0, 2^18 beam electrons at
gamma 1000 in the oblique plane wave a0 = 1, dz = 20/k0); a step is the RK4 push of every beam electron, its
deposit and currents, the march of all slices, sponge and real fields; metric: beam pushes per second.

--workload gradient: config 4 (128 x 128 modes over 120/k0, 512 slices of 0.5/k0, 20 steps of 10/k0, plasma
2e-3 n_c at 2 x 1 electrons per cell, a Gaussian driver whose amplitude along zeta is a 32-value piecewise-linear
a0(zeta) = 1.5 + N(0, 0.05), a witness of 2^18 electrons at gamma 100 with a 16-value piecewise-linear current
profile); the gradient of the witness's mean p_z after the 20 steps in the 48 values by the reverse of the march
(checkpoints every 5 steps), beside the loss alone and the tangent-linear run along one direction; metric:
gradients per second, FP64 form.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gem-pic-3d_amd"))

HBM_PEAK_GBS = 8000.0
TIME_EVERY = 16
LAMBDA0 = 2 * np.pi * 1e-6


def lwfa_deck(nx=256, ny=256, nz=1024, h=0.25, witness=True):
    """the laser-wakefield deck; witness: the trailing bunch of 2^20 electrons (config 2)"""
    import g3d_amd
    d = g3d_amd.Deck(LAMBDA0)
    d.c.nx = nx; d.c.Lx = 240.0
    d.y_uniform(ny, 240.0)
    d.zeta_uniform(nz, (nz - 1) * h)
    d.set(dz=20.0, z_end=4000.0, sponge=24.0, sponge_y=24.0, move_x=0, move_y=0, plasma_on=1, n0=1e-3,
          ppc_x=2, ppc_y=2, smooth_length=0.0, iterations=2)
    d.add_grid_laser(a0=2.0, waist=30.0, tau=15.0, zeta_c=min(200.0, 0.78 * (nz - 1) * h), z_focus=0.0)
    if witness:
        # peak 0.005 n_c: charge 0.005 (2 pi)^1.5 1.5^2 3 n_c/k0^3 = 2.40 pC at lambda0 = 2 pi um
        d.set(charge=2.40e-12, gamma=200.0, sigx=1.5, sigy=1.5, sigz=3.0, emitx=1e-3, emity=1e-3, espread=0.0,
              zeta_center=min(100.0, 0.39 * (nz - 1) * h), nmacro=1 << 20, seed=7)
    return d


def bunch_moments(p):
    """the beam moments of the reference's bunches.txt (src/sim_bunch_diag.cpp:9-61) for the live particles of a
    list (one group): weights w; first moments, then second moments centred on them (two passes).  Returns
    (w, <gamma>, eps_nx, eps_ny) in code units."""
    live = p[p["alive"] == 1]
    w = live["w"]
    W = w.sum()
    g = np.sqrt(1 + live["px"] ** 2 + live["py"] ** 2 + live["pz"] ** 2)
    x, y, ux, uy = live["x"], live["y"], live["px"], live["py"]
    m = {k: (w * v).sum() / W for k, v in (("x", x), ("y", y), ("ux", ux), ("uy", uy), ("g", g))}
    dx, dy, dux, duy = x - m["x"], y - m["y"], ux - m["ux"], uy - m["uy"]
    var = lambda a, b: (w * a * b).sum() / W
    ex = np.sqrt(max(0.0, var(dx, dx) * var(dux, dux) - var(dx, dux) ** 2))
    ey = np.sqrt(max(0.0, var(dy, dy) * var(duy, duy) - var(dy, duy) ** 2))
    return W, m["g"], ex, ey


def vacuum_deck(nx=64, nz=256, h=0.5, nmacro=1 << 18):
    """the vacuum deck: bunch of 11.4014 pC (2^18 electrons, peak 0.01 n_c for sigma 2, 2, 4/k0), gamma 1000,
    energy spread 1e-3, emittance 2e-4 um, seed 1; plane wave a0 = 1 at 2 mrad, p-polarised, c tau = 30/k0"""
    import g3d_amd
    d = g3d_amd.Deck(LAMBDA0)
    d.c.nx = nx; d.c.Lx = 80.0
    d.y_uniform(nx, 80.0)
    d.zeta_uniform(nz, (nz - 1) * h)
    Lw = (nz - 1) * h
    d.set(dz=20.0, z_end=1000.0, sponge=8.0, sponge_y=8.0, move_x=1, move_y=1, charge=11.4014089e-12, gamma=1000.0,
          sigx=2.0, sigy=2.0, sigz=4.0, emitx=2e-4, emity=2e-4, espread=1e-3, zeta_center=0.5 * Lw, nmacro=nmacro,
          seed=1, laser_on=1, laser_pol=1, a0=1.0, theta=2e-3, tau=30.0, start_fwhm=0.0, stop_fwhm=0.0)
    return d


# FP64 operations of one RK4 sub-step of the bunch push (k_bunch_push): four right-hand sides of about 355 each
# (the gather of six components over nine nodes and two slices 285, the plane wave 25, gamma and the derivatives
# 45) and the stage and update arithmetic (72); one sub-step per push on the vacuum deck (gamma 1000: neither the
# slice spacing nor the laser phase limits the 20/k0 step)
RK4_FLOP_PER_SUBSTEP = 4 * 355 + 72
# HBM bytes of one push: read x y zeta px py pz w, write the six coordinates
PUSH_BYTES = 104.0
FP64_PEAK_TFLOPS = 78.6


def run_vacuum(args, torch):
    import g3d_amd
    sim = g3d_amd.Sim(vacuum_deck())
    t0 = time.perf_counter()
    sim.init()
    t_init = time.perf_counter() - t0
    print("vacuum: init %.2f s" % t_init, file=sys.stderr, flush=True)
    for k in range(args.warmup):
        sim.step()
    info = sim.info()
    N = info["N"]
    torch.cuda.synchronize()
    sim.timing(["beam_push"])
    t0 = time.perf_counter()
    pushes = 0
    for k in range(args.steps):
        np_ = sim.info()["np"]
        sim.step()
        pushes += np_
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms, cnt = sim.timers()["beam_push"]
    avg_s = ms / 1000 / max(cnt, 1)
    np_ = sim.info()["np"]
    sim.timing(True)
    sim.step()
    torch.cuda.synchronize()
    fam = {k: {"ms_total": round(v[0], 3), "launches": v[1], "avg_us": round(1000 * v[0] / v[1], 2)}
           for k, v in sim.timers().items() if v[1]}
    sim.timing(False)
    return {"value": pushes / dt, "ms_step": 1000 * dt / args.steps, "N": N, "np": np_, "avg_us": avg_s * 1e6,
            "bpl": np_ * PUSH_BYTES, "flop": np_ * RK4_FLOP_PER_SUBSTEP, "fam": fam, "t_init": t_init,
            "slice_solves": N * args.steps / dt}


def cpu_baseline_vacuum(threads):
    """the reference's CPU build on the vacuum deck itself, 40 steps after init (about 5 s on 16 cores)"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refpy
    if not refpy.available():
        return None
    refpy.set_threads(threads)
    ref = refpy.RefSim(vacuum_deck())
    ref.init()
    steps = 40
    t = time.perf_counter()
    pushes = 0
    for _ in range(steps):
        pushes += ref.info()["np"]
        ref.step()
    dt = time.perf_counter() - t
    return {"value": pushes / dt, "unit": "pushes/s", "cores": threads, "kind": "reference",
            "sample": "reference CPU build (g++ -O3 -fopenmp, %d threads), the vacuum deck, %d steps: %.2f s"
                      % (threads, steps, dt)}


def vacuum_main(args, torch):
    r = run_vacuum(args, torch)
    cpu = None
    if not args.no_cpu:
        thr = args.cpu_threads or min(16, os.cpu_count() or 1)
        try:
            cpu = cpu_baseline_vacuum(thr)
        except Exception as e:
            print("cpu baseline failed: %s" % e, file=sys.stderr)
    avg_s = r["avg_us"] * 1e-6
    out = {
        "metric": "beam particle pushes/s (RK4 push with gathers and plane wave, deposit, field solve)",
        "value": r["value"], "unit": "pushes/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r["ms_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "vacuum bunch + plane wave 64x64x256, 2^18 beam electrons, FP64", "N": r["N"],
                   "np": r["np"], "slice_solves_per_s": r["slice_solves"]},
        "roofline": {"kernel": "k_bunch_push", "bound": "hbm", "achieved": r["bpl"] / avg_s / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": r["bpl"] / avg_s / 1e9 / HBM_PEAK_GBS,
                     "traffic": measured_traffic("k_bunch_push"), "bytes_per_launch": r["bpl"],
                     "avg_launch_us": r["avg_us"],
                     "fp64_valu": {"achieved_tflops": r["flop"] / avg_s / 1e12, "peak_tflops": FP64_PEAK_TFLOPS,
                                   "frac": r["flop"] / avg_s / 1e12 / FP64_PEAK_TFLOPS,
                                   "flop_per_push": RK4_FLOP_PER_SUBSTEP}},
        "cpu_baseline": cpu,
        "kernel_families_one_extra_step": r["fam"],
        "init_s": r["t_init"],
    }
    print(json.dumps(out))


def config4_deck(steps=20):
    """config 4: 128 x 128 modes over 120/k0, 512 slices of 0.5/k0, `steps` steps of 10/k0, plasma 2e-3 n_c at
    2 x 1 electrons per cell, the Gaussian driver (waist 25/k0) with a0(zeta) through 32 values 2/k0 apart (seeded
    1.5 + N(0, 0.05)), the witness of 2^18 electrons at gamma 100 (sigma_r 2/k0, 5 pC) with the current profile
    through 16 values (ones) over zeta 90 ... 120/k0"""
    import g3d_amd
    rng = np.random.default_rng(4)
    env = 1.5 + 0.05 * rng.standard_normal(32)
    Lw = 511 * 0.5
    d = g3d_amd.Deck(LAMBDA0)
    d.c.nx = 128; d.c.Lx = 120.0
    d.y_uniform(128, 120.0)
    d.zeta_uniform(512, Lw)
    d.set(dz=10.0, z_end=10.0 * steps, sponge=8.0, sponge_y=8.0, move_x=0, move_y=0, plasma_on=1, n0=2e-3,
          ppc_x=2, ppc_y=1, iterations=2, mixed=0, charge=5e-12, gamma=100.0, sigx=2.0, sigy=2.0, sigz=5.0,
          emitx=0.0, emity=0.0, espread=0.0, nmacro=1 << 18, seed=11)
    d.add_grid_laser(a0=1.0, waist=25.0, tau=1.0, zeta_c=Lw - 41.0, z_focus=0.0)
    d.set_envelope(0, list(env), Lw - 72.0, 2.0)
    d.set_profile([1.0] * 16, 90.0, 120.0)
    return d


def cpu_baseline_config4(threads, steps=2):
    """the reference's CPU build on config 4 itself for `steps` steps after init; the gradient by central
    differences costs 96 + 1 such runs of 20 steps"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refpy
    if not refpy.available():
        return None
    refpy.set_threads(threads)
    ref = refpy.RefSim(config4_deck())
    ref.init()
    t = time.perf_counter()
    for _ in range(steps):
        ref.step()
    dt = (time.perf_counter() - t) / steps
    return {"value": 1.0 / (97 * 20 * dt), "unit": "gradients/s", "cores": threads, "kind": "reference",
            "sample": "reference CPU build (g++ -O3 -fopenmp, %d threads), config 4, %d steps: %.2f s per step; the "
                      "48-value gradient by central differences = 97 runs of 20 steps" % (threads, steps, dt)}


def gradient_main(args, torch):
    import g3d_amd
    t = time.perf_counter()
    s = g3d_amd.Sim(config4_deck())
    s.init()
    for _ in range(20):
        s.step()
    t_loss = time.perf_counter() - t
    s.close()
    times = []
    for _ in range(max(1, args.steps // 4)):
        s = g3d_amd.Sim(config4_deck())
        t = time.perf_counter()
        loss, genv, gprof = s.gradient(20, every=5)
        times.append(time.perf_counter() - t)
        s.close()
    t_grad = min(times)
    rng = np.random.default_rng(11)
    de, dp = rng.standard_normal(32), rng.standard_normal(16)
    s = g3d_amd.Sim(config4_deck())
    t = time.perf_counter()
    _, dl = s.tangent(20, de, dp)
    t_tlm = time.perf_counter() - t
    s.close()
    cpu = None
    if not args.no_cpu:
        thr = args.cpu_threads or min(16, os.cpu_count() or 1)
        try:
            cpu = cpu_baseline_config4(thr)
        except Exception as e:
            print("cpu baseline failed: %s" % e, file=sys.stderr)
    out = {
        "metric": "config-4 gradients/s (48 parameters, 20 steps, reverse of the march)",
        "value": 1.0 / t_grad, "unit": "gradients/s", "n_gpus": 1, "steps": 20, "warmup": 0,
        "ms_per_step": 1e3 * t_grad / 20, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "config 4: 128x128x512, 2^18 witness, 20 steps, gradient in 32 envelope + 16 profile "
                               "values, checkpoints every 5 steps", "loss": loss,
                   "grad_norm": float(np.sqrt(genv @ genv + gprof @ gprof))},
        "roofline": None,
        "times_s": {"loss": t_loss, "gradient": t_grad, "tangent_one_direction": t_tlm,
                    "gradient_over_loss": t_grad / t_loss},
        "tangent_vs_adjoint_rel": abs(dl - float(genv @ de + gprof @ dp)) / abs(dl),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


N_IP = [14.53413, 29.60125, 47.4453, 77.4735, 97.8901, 552.06733, 667.04612]   # nitrogen, NIST [eV]
AMU = 1822.888486


def freehand_deck(Lw=615.5):
    """config 3: 2048 slices at Lw = 615.5/k0 (38 at Lw = 8, the CPU sample)"""
    import g3d_amd
    d = g3d_amd.Deck(LAMBDA0)
    d.c.nx = 512; d.c.Lx = 400.0
    d.y_table([0.0, 18.4, 30.4], [0.2, 0.2, 2.0], 200.0)
    if Lw > 300:
        d.zeta_table([0.0, 20.0, 60.0, 100.0, 200.0, 300.0], [0.1, 0.1, 0.15, 0.2, 0.3, 0.5], Lw)
    else:
        d.zeta_table([0.0, 2.0, 4.0], [0.1, 0.1, 0.5], Lw)
    Lz = float(d.zeta[-1])
    d.set(dz=5.0, z_end=1000.0, plasma_on=1, n0=1e-3, ppc_x=2, ppc_y=1, move_x=0, sponge=8.0, sponge_y=8.0,
          iterations=2)
    d.add_grid_laser(a0=2.5, waist=25.0, tau=12.0, zeta_c=Lz - 40.0 if Lz > 100 else 0.6 * Lz)
    d.add_ion(1.00794 * AMU, 1, 1, atoms=1.0, mobile=True, ionization=False)
    d.add_ion(14.0067 * AMU, 0, 7, atoms=0.01, mobile=True, ionization=True, ip_eV=N_IP)
    return d


def freehand_main(args, torch):
    forms = {"fp64": [False], "mixed": [True], "both": [False, True]}[args.precision]
    runs = {}
    for mixed in forms:
        r = run_form(args, mixed, torch, deck=freehand_deck())
        # one form's device state is about 130 GB at this size: keep what the comparison reads on the host and
        # release the device memory before the next form
        sim = r.pop("sim")
        info = sim.info()
        r["host"] = {"F": sim.array("F"), "ns": info["ns"], "nk": info["nk"], "nstrm": info["nstrm"],
                     "ions": [float(v) for v in sim.array("ion_stats1")]}
        sim.close()
        del sim
        runs["mixed" if mixed else "fp64"] = r
    head = runs.get("mixed") or runs["fp64"]
    compare = None
    if len(runs) == 2:
        h64, hmx = runs["fp64"]["host"], runs["mixed"]["host"]
        compare = slice_errors(hmx["F"], h64["F"], h64["ns"], h64["nk"])
        compare.update({"speedup_mixed_over_fp64": runs["mixed"]["value"] / runs["fp64"]["value"],
                        "fp64_value": runs["fp64"]["value"], "fp64_ms_per_step": runs["fp64"]["ms_step"],
                        "fp64_stream_us": runs["fp64"]["avg_us"],
                        "streaming_list_fp64_mixed": [h64["nstrm"], hmx["nstrm"]],
                        "nitrogen_charge_states_fp64": h64["ions"], "nitrogen_charge_states_mixed": hmx["ions"]})
    cpu = None
    if not args.no_cpu:
        thr = args.cpu_threads or min(16, os.cpu_count() or 1)
        try:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import refpy
            if refpy.available():
                refpy.set_threads(thr)
                ref = refpy.RefSim(freehand_deck(Lw=8.0))
                ref.init()
                t = time.perf_counter()
                ref.step()
                dtc = time.perf_counter() - t
                model, ncpu = cpu_model()
                # the reference counts no pushes: every list entry through every half-cell and pass
                cpu = {"value": ref.info()["nstrm"] * (ref.info()["N"] - 1) * 2 / dtc, "unit": "pushes/s", "cores": thr,
                       "kind": "reference", "host_cpus": ncpu, "cpu_model": model,
                       "sample": "reference CPU build (g++ -O3 -fopenmp, %d threads), the same transverse deck with "
                                 "%d slices, 1 step: %.2f s" % (thr, ref.info()["N"], dtc)}
        except Exception as e:
            print("cpu baseline failed: %s" % e, file=sys.stderr)
    mixed = "mixed" in runs
    out = {
        "metric": "particle pushes/s incl. deposit+gather+ADK, free-hand y; FP32-mixed vs FP64",
        "value": head["value"], "unit": "pushes/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-mixed" if mixed else "f64", "data": "synthetic",
        "config": {"workload": "free-hand y plasma 512x386x2048 with H and N (ADK), %s form"
                               % ("FP32-mixed" if mixed else "FP64"),
                   "N": head["N"], "nstrm": head["nstrm"], "pushes_per_step": head["pushes_per_step"],
                   "slice_solves_per_s": head["N"] * 1000.0 / head["ms_step"]},
        "roofline": {"kernel": head["kernel"], "bound": "hbm", "achieved": head["achieved"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": head["achieved"] / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": head["bpl"], "pushes_per_launch": head["push_per_launch"],
                     "design_bytes_per_launch": head["design_bpl"], "avg_launch_us": head["avg_us"],
                     "timed_launches": head["timed_launches"], "timed_one_in": TIME_EVERY},
        "cpu_baseline": cpu,
        "mixed_vs_fp64": compare,
        "kernel_families_one_extra_step": head["fam"],
        "init_s": head["t_init"],
    }
    print(json.dumps(out))


def stream_bytes_per_launch(nstrm, nk):
    """Algorithmic HBM bytes of one streaming half-cell pass (k_stream_halfcell).
    Per particle: read x y z zeta px py pz w q qm (80 B) + alive sp (8 B), write x y z zeta px py pz q qm (72 B)
    + alive (4 B) = 164 B.  Per point: gather of F slices n, n+1 (6 floats each, 48 B) and the deposit into the
    5 real planes of jsr (40 B) = 88 B."""
    return nstrm * 164.0 + nk * 88.0


def cpu_model():
    """the host CPU's model name and logical CPU count"""
    name = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return name, os.cpu_count()


def cpu_baseline(threads, witness=True):
    """The reference's own CPU build (oracle/_ref) on a bounded sample of the same workload: the full 256 x 256
    transverse grid, 262144 streaming electrons per slice and the 2^20-electron witness, 256 slices instead of 1024
    (the per-half-cell work does not depend on the window length), 2 steps after init: about 15 s on 16 cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refpy
    if not refpy.available():
        return None
    refpy.set_threads(threads)
    nz = 256
    d = lwfa_deck(nz=nz, witness=witness)
    ref = refpy.RefSim(d)
    ref.init()
    t = time.perf_counter()
    steps = 2
    pushes = 0
    for _ in range(steps):
        pushes += ref.info()["np"]
        ref.step()
        pushes += ref.info()["nstrm"] * (nz - 1) * 2
    dt = time.perf_counter() - t
    model, ncpu = cpu_model()
    return {"value": pushes / dt, "unit": "pushes/s", "cores": threads, "kind": "reference",
            "host_cpu": model, "host_logical_cpus": ncpu,
            "sample": "reference CPU build (g++ -O3 -fopenmp, %d threads), 256x256 modes, %d streaming electrons per "
                      "slice, %d slices, 2 passes, %s witness, %d steps: %.2f s"
                      % (threads, ref.info()["nstrm"], nz, "2^20" if witness else "no", steps, dt)}


def stream_bytes_mixed(nstrm, nk):
    """Design HBM bytes of one launch of k_stream_mixed (what the kernel reads and writes): per particle 144 B (read
    x y zeta z px py pz w q qm + alive, write x y zeta z px py pz + alive; the state stays FP64 in HBM, q and qm are
    constant for its electrons), per point the F gather of slices n, n+1 (48 B) and the 5 float current planes
    (20 B)."""
    return nstrm * 144.0 + nk * 68.0


# Algorithmic HBM bytes of the streaming half-cell pass (the compulsory traffic of the march): per push of a
# streaming particle (2 passes restored from the start state: the state read twice and written once, 8 words and a
# 4-byte state word) 98 B in FP64 and 64 B in FP32-mixed (z and zeta FP64); per grid point the gather of F of the
# slice (6 floats, slice n + 1 stays in the caches) 24 B and the deposit of j_x, j_y, j_z 24 B (FP64) or 12 B (FP32)
ALGO_PUSH_B = {"fp64": 98.0, "mixed": 64.0}
ALGO_POINT_B = {"fp64": 48.0, "mixed": 36.0}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "stream_traffic.json")


def measured_traffic(kernel):
    """HBM bytes per launch of the kernel from the committed byte-counter profile (FETCH_SIZE and WRITE_SIZE runs of
    rocprofv3, scaled by a plain copy of known size), or None"""
    try:
        t = json.load(open(TRAFFIC_FILE))
    except (OSError, ValueError):
        return None
    # the entry of the kernel itself or of its one template instance in the profile (k_stream_mixed<4, false>)
    hits = [v for k, v in t.items() if k == kernel or k.startswith(kernel + "<")]
    return hits[0]["bytes_per_launch"] if len(hits) == 1 else None


def run_form(args, mixed, torch, deck=None):
    import g3d_amd
    d = deck if deck is not None else lwfa_deck(nz=args.nz, witness=args.witness)
    d.set(mixed=int(mixed))
    sim = g3d_amd.Sim(d)
    t0 = time.perf_counter()
    sim.init()
    t_init = time.perf_counter() - t0
    tag = "mixed" if mixed else "fp64"
    print("%s: init %.2f s" % (tag, t_init), file=sys.stderr, flush=True)
    for k in range(args.warmup):
        sim.step()
        print("%s: warmup step %d done" % (tag, k), file=sys.stderr, flush=True)
    info = sim.info()
    N, nk = info["N"], info["nk"]
    torch.cuda.synchronize()
    # timed region: HIP events carried by one launch in TIME_EVERY of the dominant kernel (the events of a launch
    # cost a few microseconds of launch spacing; a uniform sample over the slices keeps that out of the step time)
    sim.timing(["stream_halfcell"])
    sim.timing_every(TIME_EVERY)
    i0 = sim.info()
    t0 = time.perf_counter()
    for k in range(args.steps):
        sim.step()
        print("%s: step %d done" % (tag, k), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    i1 = sim.info()
    ps, pb = i1["pushes_stream"] - i0["pushes_stream"], i1["pushes_beam"] - i0["pushes_beam"]
    pushes = ps + pb
    ms, cnt = sim.timers()["stream_halfcell"]
    avg_s = ms / 1000 / max(cnt, 1)
    sim.timing_every(1)
    nstrm = sim.info()["nstrm"]
    # the mixed stream kernel runs for Fourier y without ions; with ions or the free-hand y grid the mixed form
    # streams with the FP64 kernel (FP64 particle state, float current planes)
    c = d.c
    mixed_kernel = mixed and c.y_fourier and not any(c.ions[k].atoms > 0 for k in range(c.n_ions))
    form = "mixed" if mixed_kernel else "fp64"
    design_bpl = stream_bytes_mixed(nstrm, nk) if mixed_kernel else stream_bytes_per_launch(nstrm, nk)
    # algorithmic bytes of one stream launch: the live pushes of an average launch and the grid points
    push_per_launch = ps / (args.steps * (N - 1) * sim_passes(sim))
    bpl = push_per_launch * ALGO_PUSH_B[form] + nk * ALGO_POINT_B[form]
    kernel = "k_stream_mixed" if mixed_kernel else "k_stream_halfcell"
    # one more step, outside the timed region, with every kernel family timed (the per-kernel breakdown)
    sim.timing(True)
    sim.step()
    torch.cuda.synchronize()
    fam = {k: {"ms_total": round(v[0], 3), "launches": v[1], "avg_us": round(1000 * v[0] / v[1], 2)}
           for k, v in sim.timers().items() if v[1]}
    sim.timing(False)
    return {"sim": sim, "value": pushes / dt, "ms_step": 1000 * dt / args.steps, "N": N, "nstrm": nstrm,
            "pushes_per_step": pushes // args.steps, "stream_pushes_per_step": ps // args.steps,
            "beam_pushes_per_step": pb // args.steps, "kernel": kernel, "bpl": bpl, "design_bpl": design_bpl,
            "push_per_launch": push_per_launch, "avg_us": avg_s * 1e6, "achieved": bpl / avg_s / 1e9,
            "traffic": measured_traffic(kernel), "fam": fam, "t_init": t_init, "timed_launches": cnt}


def sim_passes(sim):
    return int(sim._deck.c.iterations)


def copy_calibration(mib):
    """dst <- src over mib MiB, 5 times, with torch's copy kernel: each dispatch reads and writes mib MiB"""
    import torch
    n = int(mib * (1 << 20)) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_each_way": n * 4, "dispatches": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nz", type=int, default=1024)
    ap.add_argument("--no-witness", dest="witness", action="store_false",
                    help="the laser-wakefield deck without the witness bunch (config 1)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--precision", choices=("both", "mixed", "fp64"), default="both",
                    help="both (default): the FP64 and the FP32-mixed march on the same deck; the line reports the "
                         "mixed one and its field error against FP64")
    ap.add_argument("--workload", choices=("lwfa", "vacuum", "freehand", "gradient"), default="lwfa",
                    help="lwfa (default): the plasma march of the headline metric; vacuum: the beam-push deck; "
                         "freehand: config 3 (free-hand y, ions, ADK)")
    ap.add_argument("--copy-calib", type=float, default=0.0,
                    help="only copy a buffer of this many MiB a few times (the reference kernel that scales the HBM "
                         "byte counters of a profile) and exit")
    args = ap.parse_args()
    if args.copy_calib > 0:
        copy_calibration(args.copy_calib)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU (no collectives); --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "no GPU"
    if args.workload == "vacuum":
        vacuum_main(args, torch)
        return
    if args.workload == "freehand":
        freehand_main(args, torch)
        return
    if args.workload == "gradient":
        gradient_main(args, torch)
        return

    forms = {"fp64": [False], "mixed": [True], "both": [False, True]}[args.precision]
    runs = {}
    for mixed in forms:
        runs["mixed" if mixed else "fp64"] = run_form(args, mixed, torch)
    head = runs.get("mixed") or runs["fp64"]
    compare = None
    if len(runs) == 2:
        compare = compare_forms(runs["fp64"], runs["mixed"])
        compare.update({"speedup_mixed_over_fp64": runs["mixed"]["value"] / runs["fp64"]["value"],
                        "fp64_value": runs["fp64"]["value"], "fp64_ms_per_step": runs["fp64"]["ms_step"],
                        "fp64_stream_us": runs["fp64"]["avg_us"]})
    cpu = None
    if not args.no_cpu:
        thr = args.cpu_threads or min(16, os.cpu_count() or 1)
        try:
            cpu = cpu_baseline(thr, args.witness)
        except Exception as e:  # the baseline is reported, never the target
            print("cpu baseline failed: %s" % e, file=sys.stderr)
    mixed = "mixed" in runs
    out = {
        "metric": "particle pushes/s incl. deposit+gather; FP32-mixed vs FP64 at <=1e-4 field err",
        "value": head["value"], "unit": "pushes/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-mixed" if mixed else "f64", "data": "synthetic",
        "config": {"workload": "LWFA plasma march 256x256x%d, 262144 streaming e- per slice, 2 passes, %s"
                               "witness 2^20 e- at gamma 200, %s form"
                               % (args.nz, "" if args.witness else "no ", "FP32-mixed" if mixed else "FP64"),
                   "N": head["N"], "nstrm": head["nstrm"], "pushes_per_step": head["pushes_per_step"],
                   "stream_pushes_per_step": head["stream_pushes_per_step"],
                   "witness_pushes_per_step": head["beam_pushes_per_step"],
                   # a slice solve: one slice of the march for all modes with its transforms (every pass of every
                   # half-cell, and the front slice)
                   "slice_solves_per_s": ((head["N"] - 1) * sim_passes(head["sim"]) + 1) * 1000.0 / head["ms_step"]},
        "roofline": {"kernel": head["kernel"], "bound": "hbm", "achieved": head["achieved"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": head["achieved"] / HBM_PEAK_GBS, "traffic": head["traffic"],
                     "bytes_per_launch": head["bpl"], "pushes_per_launch": head["push_per_launch"],
                     "design_bytes_per_launch": head["design_bpl"], "avg_launch_us": head["avg_us"],
                     "timed_launches": head["timed_launches"], "timed_one_in": TIME_EVERY},
        "cpu_baseline": cpu,
        "mixed_vs_fp64": compare,
        "kernel_families_one_extra_step": head["fam"],
        "init_s": head["t_init"],
    }
    print(json.dumps(out))


def slice_errors(Fmx, F64, ns, nk):
    """the real fields F of every slice of two runs (E and B of the slice together), relative L2 against FP64"""
    A, B = Fmx.reshape(6, ns + 2, nk), F64.reshape(6, ns + 2, nk)
    nb, de = np.zeros(ns), np.zeros(ns)
    for m0 in range(0, ns, 64):                      # in blocks of slices (F of config 3 is 10 GB)
        a = A[:, m0:min(m0 + 64, ns)].astype(np.float64)
        b = B[:, m0:min(m0 + 64, ns)].astype(np.float64)
        nb[m0:m0 + a.shape[1]] = np.sqrt((b ** 2).sum(axis=(0, 2)))
        de[m0:m0 + a.shape[1]] = np.sqrt(((a - b) ** 2).sum(axis=(0, 2)))
    err = de / np.maximum(nb, 1e-300)
    worst = int(err.argmax())
    big = nb >= 1e-2 * nb.max()          # slices carrying at least 1 % of the largest slice's field
    out = {"F_rel_l2_max_over_slices": float(err.max()), "worst_slice": worst,
           "worst_slice_norm_over_max": float(nb[worst] / nb.max()),
           "F_rel_l2_p50_p99_over_slices": [float(v) for v in np.percentile(err, [50, 99])],
           "F_rel_l2_all_slices": float(np.sqrt((de ** 2).sum() / (nb ** 2).sum())),
           "F_rel_l2_max_over_slices_above_1pct": float(err[big].max()), "slices": int(len(nb))}
    return out


def compare_forms(r64, rmx):
    """the two runs took the same steps from the same start: the real fields F of every slice (relative L2 against
    FP64, every slice), and the witness moments of the reference's bunches.txt"""
    s64, smx = r64["sim"], rmx["sim"]
    info = s64.info()
    out = slice_errors(smx.array("F"), s64.array("F"), info["ns"], info["nk"])
    p64, pmx = s64.array("parts"), smx.array("parts")
    if len(p64):
        g0 = s64._deck.c.gamma
        W64, g64, ex64, ey64 = bunch_moments(p64)
        Wmx, gmx, exmx, eymx = bunch_moments(pmx)
        out.update({"witness_gain_fp64": g64 - g0, "witness_gain_mixed": gmx - g0,
                    "witness_gain_rel_err": abs((gmx - g0) - (g64 - g0)) / abs(g64 - g0),
                    "witness_emit_nx_fp64": ex64, "witness_emit_nx_rel_err": abs(exmx - ex64) / ex64,
                    "witness_emit_ny_rel_err": abs(eymx - ey64) / ey64,
                    "witness_live": [int((p64["alive"] == 1).sum()), int((pmx["alive"] == 1).sum())]})
    return out


if __name__ == "__main__":
    main()
