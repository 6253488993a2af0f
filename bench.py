"""Benchmark: energy + force atom-steps/s of the HDNNP engine on one MI355X.

Workload (default): water-like H/O box of 196,608 atoms (65,536 molecules,
cubic 124.2 A cell), 8 radial G2 per neighbour element + 8 angular G3 per
element pair, rc = 6.35 A cosine cutoff, min/max scaling from this
structure's SF statistics (computed on the GPU, the nnp-scaling
definition), random-init nsf-25-25-1 tanh networks (Xavier-uniform, seeded).
One step = one full Mode::evaluateNNP on the GPU: neighbour list rebuilt,
symmetry functions, networks, energy, forces; inputs resident in HBM.
Both paths are timed in the same run: the plain FP64 path and the mixed-
precision path (the reported value), with the mixed path's force and
energy error against the FP64 results.

Prints one JSON line (see the driver contract) with a roofline object for
the dominant kernel and a CPU baseline (the oracle, OpenMP on the host
cores, on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from n2p2_amd import costmodel, engine, workloads as W  # noqa: E402
from n2p2_amd import surrogate as S  # noqa: E402

PEAK_FP64_TFLOPS = 78.6     # MI355X vector FP64
PEAK_FP32_TFLOPS = 157.3    # MI355X vector FP32 (published; packed math)
PEAK_HBM_GBS = 8000.0       # MI355X HBM3E


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_model(tmp, r, elem, box):
    """Model files for the water workload; scaling from GPU statistics."""
    hidden = (25, 25)
    W.write_model(tmp, W.water_settings(), W.WATER_ELEMENTS, (1, 8), hidden,
                  seed=7)
    m0 = engine.Model.load_dir(tmp, scaling=False)
    c0 = engine.Context(m0)
    c0.set_structure(r, elem, box)
    c0.calculate_neighbor_list(m0.max_cutoff)
    stats = []
    for e in range(m0.num_elements):
        _, st = c0.sf_statistics(e)
        st = st.copy()
        flat = st[:, 1] - st[:, 0] <= 0
        st[flat, 1] = st[flat, 0] + 1.0
        stats.append(st)
    c0.close()
    W.write_scaling(os.path.join(tmp, "scaling.data"), stats)
    return engine.Model.load_dir(tmp), hidden


def cpu_baseline(tmp, seconds_target=20.0):
    """The oracle (CPU restatement of the reference path, OpenMP over all
    host cores) on a 24,576-atom sample of the same water workload: cell-list
    neighbour list + SFs with derivatives + networks + energy + forces,
    median of 3 timed steps after one warm-up."""
    from tests.oracle_lib import OModel, OStruct, lib
    nmol, L = 8192, 62.1
    r, elem, box = W.water(nmol, L, 4321)
    m = OModel.from_dir(tmp)
    s = OStruct(m, r, elem, box)
    s.evaluate(cells=True)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        s.evaluate(cells=True)
        times.append(time.perf_counter() - t0)
        if sum(times) > seconds_target:
            break
    t = float(np.median(times))
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": len(r) / t, "unit": "atom-steps/s",
            "cores": int(lib().orc_num_threads()), "kind": "port",
            "sample": "oracle on a 24,576-atom water box of the same model "
                      "(8192 molecules, 62.1 A), cell-list neighbour list, "
                      "%d timed steps, median %.3f s/step, %s"
                      % (len(times), t, cpu)}


def run_path(ctx, precision, steps, warmup):
    """Times `steps` full evaluations of one precision path after `warmup`
    untimed ones; returns the step time, energy, forces and per-kernel
    device times (HIP events on the engine's stream)."""
    ctx.set_precision(precision)
    spec = ctx.specialized_active()
    for _ in range(warmup):
        ctx.evaluate(engine.FORCES_REFERENCE)
    # evaluate() returns after its stream has drained (the energy and the
    # forces are complete), so each side of the timed region is synchronous.
    # The per-kernel device times are HIP events recorded on the engine's
    # stream around each kernel class during the same timed steps.
    ctx.profile(True)
    ctx.profile_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        energy = ctx.evaluate(engine.FORCES_REFERENCE)
    dt = (time.perf_counter() - t0) / steps
    kt = {k: ctx.profile_avg_ms(k)[0]
          for k in ("neighbor", "partner", "symfunc", "network", "force",
                    "gather")}
    ctx.profile(False)
    forces = ctx.forces()
    names = {"force": ctx.kernel_name("force"),
             "symfunc": ctx.kernel_name("sf"), "network": "k_nn_mfma"}
    return dt, energy, forces, kt, names if spec else None


def run_host_io(ctx, r, precision, steps, warmup):
    """The same evaluation with the caller's host arrays on both sides of
    every step: positions copied in (hdnnp_set_positions) and forces copied
    out (hdnnp_get_forces), as a caller without device buffers pays."""
    ctx.set_precision(precision)
    rr = np.ascontiguousarray(r, dtype=np.float64)
    for _ in range(warmup):
        ctx.set_positions(rr)
        ctx.evaluate(engine.FORCES_REFERENCE)
        ctx.forces()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.set_positions(rr)
        ctx.evaluate(engine.FORCES_REFERENCE)
        f = ctx.forces()
    dt = (time.perf_counter() - t0) / steps
    return dt, f


def tune(ctx, precision, log, label):
    """Sweeps the launch forms (member chunk, waves-per-SIMD bounds of the
    specialised SF and force kernels, network tiles per wave) on the loaded
    box and applies the fastest; untimed.  Returns the choice, the sweep
    time and the rows (columns Context.TUNE_COLUMNS)."""
    ctx.set_precision(precision)
    t0 = time.perf_counter()
    rows = ctx.autotune(engine.FORCES_REFERENCE, sweep=True, reps=10)
    secs = time.perf_counter() - t0
    log("tuning %s: %d settings in %.1f s, choice (chunk, sf waves, force "
        "waves, nn tiles) = %s" % (label, len(rows), secs, ctx.tuning()))
    log("  " + " ".join("%11s" % c for c in engine.Context.TUNE_COLUMNS))
    for row in rows:
        log("  " + " ".join("%11.4f" % x for x in row))
    return {"choice": dict(zip(("member_chunk", "sf_waves_per_simd",
                                "force_waves_per_simd", "nn_tiles_per_wave"),
                               ctx.tuning())),
            "sweep_seconds": secs, "settings_timed": len(rows),
            "rows": [[round(float(x), 4) for x in row] for row in rows]}


def run_batch(tmp, steps, warmup, log, sweep=True):
    """The 512-structure batch (three elements B/C/N at 0.4/0.4/0.2, 16-512
    atoms each, 0.08 atoms/A^3, 10 G2 per element and 12 G9 per element
    pair, polynomial cutoff 6 A, 102-30-30-1 softplus): one E+F evaluation
    of all structures per step, neighbour lists rebuilt, FP64 and mixed;
    the mixed path's force error against FP64 over all atoms.  With
    `sweep`, the FP64 and mixed legs first sweep the launch forms on the
    batch (untimed); the FP16 / BF16 legs keep the mixed choice."""
    sts = W.batch_structures(512, 16, 512, seed=2024)
    d = os.path.join(tmp, "batch_model")
    r0, e0, b0 = sts[0]
    S.write_model_gpu_scaled(d, W.mixed_settings(), W.MIXED_ELEMENTS,
                             (5, 6, 7), (30, 30), r0, e0, b0, seed=11)
    model = engine.Model.load_dir(d)
    ctx = engine.Context(model)
    ctx.set_structures(sts)
    out = {"structures": len(sts), "atoms": int(ctx.n)}
    res = {}
    for name, prec in (("fp64", engine.PRECISION_FP64),
                       ("mixed", engine.PRECISION_MIXED),
                       ("mixed16", engine.PRECISION_MIXED16),
                       ("mixed_bf16", engine.PRECISION_MIXED_BF16)):
        ctx.set_precision(prec)
        if sweep and name in ("fp64", "mixed"):
            rows = ctx.autotune(engine.FORCES_REFERENCE, sweep=True, reps=5)
            log("batch tuning %s: %d settings, choice %s"
                % (name, len(rows), ctx.tuning()))
        for _ in range(warmup):
            ctx.evaluate(engine.FORCES_REFERENCE)
        ctx.profile(True)
        ctx.profile_reset()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.evaluate(engine.FORCES_REFERENCE)
        dt = (time.perf_counter() - t0) / steps
        kt = {k: ctx.profile_avg_ms(k)[0]
              for k in ("neighbor", "partner", "symfunc", "network",
                        "force", "gather")}
        ctx.profile(False)
        res[name] = (ctx.forces(), ctx.structure_energies())
        out[name] = {"atom_steps_per_s": ctx.n / dt, "ms_per_step": dt * 1e3,
                     "kernel_ms": kt,
                     "tuning": dict(zip(("member_chunk", "sf_waves_per_simd",
                                         "force_waves_per_simd",
                                         "nn_tiles_per_wave"),
                                        ctx.tuning()))}
    for name in ("mixed", "mixed16", "mixed_bf16"):
        df = res[name][0] - res["fp64"][0]
        de = np.abs(res[name][1] - res["fp64"][1]) / np.abs(res["fp64"][1])
        out[name]["error_vs_fp64"] = {
            "force_rmse_eV_A": float(np.sqrt(np.mean(df ** 2))),
            "force_max_eV_A": float(np.abs(df).max()),
            "structure_energy_rel_max": float(de.max())}
    out["mixed_vs_plain_fp64"] = (out["mixed"]["atom_steps_per_s"]
                                  / out["fp64"]["atom_steps_per_s"])
    ctx.close()
    log("batch of 512 structures: %s" % out)
    return out


def run_md(model, r, elem, box, precision, steps, warmup, skin=1.0,
           dt_fs=0.25, temp=300.0, tuned=False):
    """NVE MD of the same box (velocity Verlet, Verlet-skin list, HIP-graph
    steps): `warmup` untimed steps, then `steps` timed ones.  `tuned`: a
    launch form (chunk, SF waves, force waves, tiles), True for the choice
    a sweep of this process cached, False for the built-in one.  md_run
    returns after its stream has drained, so the timed region is
    synchronous on both sides."""
    ctx = engine.Context(model)
    ctx.set_precision(precision)
    ctx.set_structure(r, elem, box)
    if isinstance(tuned, tuple):
        ctx.set_tuning(*tuned)
    elif tuned:
        ctx.autotune()   # the choice cached by the sweep of this process
    masses = np.array(W.WATER_MASSES)
    v = W.maxwell_boltzmann(masses[elem], temp, seed=20261202)
    ctx.md_init(masses, dt_fs, skin=skin, v=v)
    ctx.md_run(warmup)
    reb0 = ctx.md_rebuilds()
    t0 = time.perf_counter()
    ep, ek = ctx.md_run(steps)
    dt = (time.perf_counter() - t0) / steps
    et = ep + ek
    out = {"atom_steps_per_s": len(r) / dt, "ms_per_step": dt * 1e3,
           "steps": steps, "rebuilds_in_timed_steps": ctx.md_rebuilds() - reb0,
           "energy_spread_eV_per_atom": float((et.max() - et.min()) / len(r))}
    ctx.close()
    return out


def train_surrogate(tmp, log):
    """The distilled radial surrogate (8 G2 per element, 8-16-16-1
    softplus), trained with the engine's kernels on 64 frames of the full
    model's own 300 K MD of a 1536-atom box of the same water generator
    (20 epochs, Adam lr 2e-3); returns the model and the loss history."""
    rs, es, bs = W.water(512, 24.84, 77)
    full = engine.Model.load_dir(tmp)
    sd = os.path.join(tmp, "surrogate")
    S.write_model_gpu_scaled(sd, S.surrogate_settings(), W.WATER_ELEMENTS,
                             (1, 8), S.SURROGATE_HIDDEN, rs, es, bs, seed=21)
    masses = np.array(W.WATER_MASSES)
    R, E, F = S.generate_frames(full, rs, es, bs, masses, nframes=64,
                                stride=8)
    sm = engine.Model.load_dir(sd)
    ctx = engine.Context(sm)
    t0 = time.perf_counter()
    hist = S.train(ctx, es, bs, R, E, F, epochs=20, lr=2e-3)
    ttrain = time.perf_counter() - t0
    S.write_weights(sd, ctx.get_weights(),
                    [sm.num_sf(k) for k in range(sm.num_elements)], (1, 8))
    ctx.close()
    log("surrogate trained in %.2f s: loss %s -> %s" % (ttrain, hist[0],
                                                        hist[-1]))
    return engine.Model.load_dir(sd), hist, ttrain


def _drift(etot, n):
    """NVE energy drift of a run: slope of a least-squares line through the
    total energy per atom (eV per atom per step) and the largest excursion
    from the first value (eV per atom)."""
    et = np.asarray(etot) / n
    k = np.arange(len(et))
    slope = float(np.polyfit(k, et, 1)[0]) if len(et) > 1 else 0.0
    return {"slope_eV_per_atom_per_step": slope,
            "max_excursion_eV_per_atom": float(np.abs(et - et[0]).max())}


def run_c4(model, r, elem, box, tmp, log, nframes=2000, stride=2,
           epochs=20, nouter=2000, gate=5e-3, chunk=100, equil=2000,
           nouter_extra=500):
    """The learned surrogate for multiple time steps on the 196,608-atom
    box: frames of the full model's own 300 K trajectory (mixed path, dt
    0.25 fs, `equil` steps of equilibration, velocities rescaled to 300 K
    every 20 steps, a frame every `stride` steps), the distilled radial
    surrogate (4 eta
    x 2 neighbour elements G2, 8-16-16-1 softplus) trained on them with the
    engine's gradient and Adam kernels (lr 1e-3, `epochs` epochs, one frame
    per step), then, from the trajectory's last state, `nouter` outer
    steps of 1 fs (4 surrogate substeps of 0.25 fs) of r-RESPA: with the
    force-RMSE gate, without it, and the full model's plain velocity Verlet
    at 0.25 fs over the same time; then `nouter_extra` outer steps without
    the gate for the untrained surrogate and with eager steps instead of
    HIP graphs.  Throughput counts 0.25 fs steps."""
    masses = np.array(W.WATER_MASSES)
    n = len(r)
    out = {"frames": nframes, "frame_stride_steps": stride, "epochs": epochs,
           "outer_steps": nouter, "gate_eV_A": gate}
    t0 = time.perf_counter()
    R, E, F, last = S.generate_frames(model, r, elem, box, masses, nframes,
                                      stride=stride, dt=0.25, temp=300.0,
                                      seed=4242,
                                      precision=engine.PRECISION_MIXED,
                                      final_state=True, equil=equil,
                                      rescale_every=20, log=log)
    out["frame_generation_s"] = time.perf_counter() - t0
    log("C4: %d frames in %.1f s" % (nframes, out["frame_generation_s"]))
    sd = os.path.join(tmp, "surrogate_c4")
    S.write_model_gpu_scaled(sd, S.surrogate_settings(), W.WATER_ELEMENTS,
                             (1, 8), S.SURROGATE_HIDDEN, R[0], elem, box,
                             seed=21)
    untrained = engine.Model.load_dir(sd)
    ctx = engine.Context(untrained)
    t0 = time.perf_counter()
    hist = S.train(ctx, elem, box, R, E, F, epochs=epochs, lr=1e-3, log=log)
    out["train_s"] = time.perf_counter() - t0
    out["train_loss_first_epoch"] = list(hist[0])
    out["train_loss_last_epoch"] = list(hist[-1])
    td = os.path.join(tmp, "surrogate_c4_trained")
    os.makedirs(td, exist_ok=True)
    for f in ("input.nn", "scaling.data"):
        with open(os.path.join(sd, f)) as a, open(os.path.join(td, f), "w") as b:
            b.write(a.read())
    S.write_weights(td, ctx.get_weights(),
                    [untrained.num_sf(k) for k in range(untrained.num_elements)],
                    (1, 8))
    ctx.close()
    del R, F
    trained = engine.Model.load_dir(td)
    # The runs below continue the frames' trajectory from its last state
    # (NVE from there).  The generated box starts far from equilibrium
    # (lattice oxygens, random orientations): without the thermostat of
    # the equilibration and the frames it heats up to thousands of kelvin.
    r, v = last
    out["start_temperature_K"] = S.temperature(v, masses[elem])

    def mts(sur_model, g, label, graph=True, nouter=nouter):
        full = engine.Context(model)
        sur = engine.Context(sur_model)
        full.set_precision(engine.PRECISION_MIXED)
        for c in (full, sur):
            c.set_structure(r, elem, box)
        sur.mts_init(full, masses, dt_outer=1.0, n_inner=4, skin=1.0, v=v,
                     rmse_gate=g, graph=graph)
        # In chunks of outer steps (the run continues across calls), with a
        # progress line per chunk.
        parts = []
        t1 = time.perf_counter()
        try:
            done = 0
            while done < nouter:
                k = min(chunk, nouter - done)
                parts.append(sur.mts_run(k))
                done += k
                log("C4 %s: %d/%d outer steps, %.1f s, rebuilds full %d "
                    "surrogate %d, fallbacks %d"
                    % (label, done, nouter, time.perf_counter() - t1,
                       full.md_rebuilds(), sur.md_rebuilds(),
                       int(sum(int(np.sum(p[3] == 1)) for p in parts))))
        except RuntimeError as err:
            full.close()
            sur.close()
            log("C4 %s: %s" % (label, err))
            return {"failed": str(err).strip()}
        dt = time.perf_counter() - t1
        ep, ek, rm, md = (np.concatenate([p[q] for p in parts])
                          for q in range(4))
        res = {"atom_steps_per_s": n * 4 * nouter / dt,
               "ms_per_outer_step": dt / nouter * 1e3,
               "fallback_fraction": float(np.mean(md == 1)),
               "surrogate_force_rmse_eV_A": {
                   "mean": float(np.mean(rm)), "max": float(np.max(rm))},
               "drift": _drift(ep + ek, n)}
        full.close()
        sur.close()
        log("C4 %s: %s" % (label, res))
        return res

    out["mts_gated"] = mts(trained, gate, "MTS, gate %g eV/A" % gate)
    out["mts_no_gate"] = mts(trained, 1e30, "MTS without gate")
    out["mts_no_gate_untrained"] = mts(untrained, 1e30,
                                       "MTS without gate, untrained",
                                       nouter=nouter_extra)
    # The trained run with eager steps instead of HIP graphs: the graphs
    # are captured again after every list rebuild.
    out["mts_no_gate_eager"] = mts(trained, 1e30, "MTS without gate, eager",
                                   graph=False, nouter=nouter_extra)
    ctx = engine.Context(model)
    ctx.set_precision(engine.PRECISION_MIXED)
    ctx.set_structure(r, elem, box)
    ctx.md_init(masses, 0.25, skin=1.0, v=v)
    t1 = time.perf_counter()
    parts = []
    done = 0
    while done < 4 * nouter:
        k = min(4 * chunk, 4 * nouter - done)
        parts.append(ctx.md_run(k))
        done += k
        log("C4 full model: %d/%d steps, %.1f s"
            % (done, 4 * nouter, time.perf_counter() - t1))
    dt = time.perf_counter() - t1
    ep = np.concatenate([p[0] for p in parts])
    ek = np.concatenate([p[1] for p in parts])
    out["full_verlet_0.25fs"] = {"atom_steps_per_s": n * 4 * nouter / dt,
                                 "drift": _drift((ep + ek)[3::4], n)}
    ctx.close()
    log("C4 full model: %s" % out["full_verlet_0.25fs"])
    for k in ("mts_gated", "mts_no_gate", "mts_no_gate_eager"):
        if "failed" in out[k]:
            continue
        out[k]["speedup_vs_full"] = (out[k]["atom_steps_per_s"]
                                     / out["full_verlet_0.25fs"][
                                         "atom_steps_per_s"])
    return out


PEAK_MFMA_FP32_TFLOPS = 157.3  # MI355X dense FP32 matrix
PEAK_MFMA_FP64_TFLOPS = 78.6


def other_rooflines(ktm, kt64, counts, nnf, n):
    """The neighbour build against HBM bandwidth (algorithmic bytes:
    positions read, index, image key, dr, d and partner written per entry)
    and the network layers against the MFMA peak of their operand type
    (algorithmic flops: forward and backward GEMMs), both from the HIP-event
    kernel times of the timed steps.  The partner map runs on a second
    stream beside the SF kernel; its time is added to the list kernels'
    here (kernel time, not wall time)."""
    nb = costmodel.neighbor_bytes(n, int(counts[0]))
    out = {}
    for name, kt in (("mixed", ktm), ("plain_fp64", kt64)):
        ms_nl = kt["neighbor"] + kt.get("partner", 0.0)
        ms_nn = kt["network"]
        peak = PEAK_MFMA_FP32_TFLOPS if name == "mixed" \
            else PEAK_MFMA_FP64_TFLOPS
        a_nn = nnf / (ms_nn * 1e-3) / 1e12
        a_nl = nb / (ms_nl * 1e-3) / 1e9
        out[name] = {
            "neighbor_build": {"bound": "hbm", "achieved": a_nl,
                               "peak": PEAK_HBM_GBS, "unit": "GB/s",
                               "frac": a_nl / PEAK_HBM_GBS,
                               "algorithmic_bytes": nb, "ms": ms_nl},
            "networks": {"bound": "mfma", "achieved": a_nn, "peak": peak,
                         "unit": "TFLOP/s", "frac": a_nn / peak,
                         "algorithmic_flops": nnf, "ms": ms_nn}}
    return out


def roofline(kt, counts, nnf, n, fp32, spec):
    """Roofline object for the dominant kernel of one path; spec: the names
    of the model-specialised SF and force kernels that ran (None: the
    generic kernels ran)."""
    fwd = int(costmodel.forward_flops(counts))
    bwd = int(costmodel.backward_flops(counts))
    peak = PEAK_FP32_TFLOPS if fp32 else PEAK_FP64_TFLOPS
    dom = max(kt, key=kt.get)
    flops = {"force": bwd, "symfunc": fwd, "network": nnf}.get(dom)
    ms = kt[dom]
    if flops is None:
        nb = costmodel.neighbor_bytes(n, int(counts[0]))
        achieved = nb / (ms * 1e-3) / 1e9
        return {"kernel": "neighbor build", "bound": "hbm",
                "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS, "traffic": None,
                "avg_launch_ms": ms}
    if spec:
        kname = spec[dom]
    else:
        kname = {"force": "k_force_backward", "symfunc": "k_sf_forward",
                 "network": "k_nn_mfma"}[dom]
    achieved = flops / (ms * 1e-3) / 1e12
    traffic = None
    tp = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tp) and n == 196608:
        with open(tp) as f:
            tk = json.load(f).get("kernels", {})
        # The waves-per-SIMD bound (the _w suffix) does not change what a
        # kernel reads and writes: a profile of another bound of the same
        # arithmetic and member chunk stands in when the sweep chose a
        # bound that was not profiled.
        base = kname.rsplit("_w", 1)[0] + "_w"
        hit = tk.get(kname) or next(
            (v for k, v in sorted(tk.items()) if k.startswith(base)), None)
        if hit:
            traffic = hit["read_bytes"] + hit["write_bytes"]
    return {"kernel": kname, "bound": "valu-fp32" if fp32 else "valu-fp64",
            "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": traffic,
            "traffic_source": ("profiles/traffic.json (rocprofv3 FETCH_SIZE "
                               "+ WRITE_SIZE, copy-kernel calibrated)")
                              if traffic else None,
            "algorithmic_flops_per_launch": flops, "avg_launch_ms": ms}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nmol", type=int, default=65536)
    ap.add_argument("--box", type=float, default=124.2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the built-in launch forms (no sweep)")
    ap.add_argument("--set-tuning", action="append", default=[],
                    metavar="PATH=C,S,F,T",
                    help="apply a launch form without a sweep (PATH fp64 "
                         "or mixed; member chunk, SF and force waves per "
                         "SIMD, network tiles per wave), e.g. for a "
                         "profiled run of the form a sweep chose")
    ap.add_argument("--c4-frames", type=int, default=2000)
    ap.add_argument("--c4-epochs", type=int, default=20)
    ap.add_argument("--c4-outer", type=int, default=2000)
    ap.add_argument("--path", choices=("both", "mixed", "fp64", "md",
                                       "batch", "mts"),
                    default="both",
                    help="time one path only (for profiling runs; the "
                         "reported line needs both)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; the multi-GPU path "
              "(spatial decomposition with halo exchange) is not built, so "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    fixed = {}
    for item in args.set_tuning:
        key, _, vals = item.partition("=")
        if key not in ("fp64", "mixed") or len(vals.split(",")) != 4:
            ap.error("--set-tuning takes fp64=C,S,F,T or mixed=C,S,F,T")
        fixed[key] = tuple(int(v) for v in vals.split(","))

    r, elem, box = W.water(args.nmol, args.box, 20261202)
    n = len(r)
    with tempfile.TemporaryDirectory() as tmp:
        log("workload: water %d atoms, L=%.2f" % (n, args.box))
        model, hidden = make_model(tmp, r, elem, box)
        ctx = engine.Context(model)
        ctx.set_structure(r, elem, box)
        ctx.calculate_neighbor_list(model.max_cutoff)
        counts = [int(x) for x in ctx.work_counts()]
        log("work counts", counts)
        nnf = int(sum(int(np.count_nonzero(elem == e))
                      * costmodel.network_flops(
                          [[model.num_sf(e)] + list(hidden) + [1]])
                      for e in range(model.num_elements)))

        if args.path == "batch":
            run_batch(tmp, args.steps, args.warmup, log,
                      sweep=not args.no_tune)
            log("batch only: no JSON line")
            return
        if args.path == "mts":
            c4 = run_c4(model, r, elem, box, tmp, log,
                        nframes=args.c4_frames, epochs=args.c4_epochs,
                        nouter=args.c4_outer)
            print(json.dumps({"c4_multiple_time_steps": c4}), flush=True)
            log("MTS only: no bench line")
            return
        if args.path == "md":
            res = run_md(model, r, elem, box, engine.PRECISION_MIXED,
                         args.steps, args.warmup,
                         tuned=fixed.get("mixed", False))
            log("MD mixed only: %s (no JSON line)" % res)
            return
        tuning = {}

        def choose(precision, label):
            if label in fixed:
                ctx.set_precision(precision)
                ctx.set_tuning(*fixed[label])
                log("tuning %s: set to %s" % (label, fixed[label]))
                tuning[label] = {"choice": dict(zip(
                    ("member_chunk", "sf_waves_per_simd",
                     "force_waves_per_simd", "nn_tiles_per_wave"),
                    fixed[label])), "sweep_seconds": 0.0,
                    "settings_timed": 0, "rows": []}
            elif not args.no_tune:
                tuning[label] = tune(ctx, precision, log, label)

        if args.path in ("both", "fp64"):
            choose(engine.PRECISION_FP64, "fp64")
            dt64, e64, f64, kt64, sp64 = run_path(
                ctx, engine.PRECISION_FP64, args.steps, args.warmup)
            log("plain FP64: %.3f ms/step, E=%.10f, kernel ms %s"
                % (dt64 * 1e3, e64, kt64))
        if args.path in ("both", "mixed"):
            choose(engine.PRECISION_MIXED, "mixed")
        if args.path != "fp64":
            dtm, em, fm, ktm, spm = run_path(
                ctx, engine.PRECISION_MIXED, args.steps, args.warmup)
            log("mixed: %.3f ms/step, E=%.10f, kernel ms %s"
                % (dtm * 1e3, em, ktm))
        if args.path != "both":
            log("single-path run: no JSON line")
            ctx.close()
            return
        # The FP16 / BF16 legs run the mixed path's SF and force kernels:
        # they keep its tuning (the context's current one).
        dt16, e16, f16, kt16, _ = run_path(ctx, engine.PRECISION_MIXED16,
                                           args.steps, args.warmup)
        dtb, eb, fb, ktb, _ = run_path(ctx, engine.PRECISION_MIXED_BF16,
                                       args.steps, args.warmup)
        dbf = fb - f64
        errbf = {"force_rmse_eV_A": float(np.sqrt(np.mean(dbf ** 2))),
                 "force_max_eV_A": float(np.abs(dbf).max()),
                 "energy_rel": float(abs(eb - e64) / abs(e64))}
        log("mixed with BF16 networks: %.3f ms/step, errors %s"
            % (dtb * 1e3, errbf))
        log("mixed with FP16 networks: %.3f ms/step, E=%.10f, kernel ms %s"
            % (dt16 * 1e3, e16, kt16))
        dth, fh = run_host_io(ctx, r, engine.PRECISION_MIXED, args.steps,
                              args.warmup)
        log("mixed with host positions in and forces out: %.3f ms/step"
            % (dth * 1e3))
        df = fm - f64
        err = {"force_rmse_eV_A": float(np.sqrt(np.mean(df ** 2))),
               "force_max_eV_A": float(np.abs(df).max()),
               "force_rms_scale_eV_A": float(np.sqrt(np.mean(f64 ** 2))),
               "energy_rel": float(abs(em - e64) / abs(e64)),
               "reference": "plain FP64 GPU path of the same run (parity-"
                            "tested against the FP64 oracle)"}
        log("mixed vs FP64 errors", err)
        d16 = f16 - f64
        err16 = {"force_rmse_eV_A": float(np.sqrt(np.mean(d16 ** 2))),
                 "force_max_eV_A": float(np.abs(d16).max()),
                 "energy_rel": float(abs(e16 - e64) / abs(e64))}
        log("mixed16 vs FP64 errors", err16)
        fwd = int(costmodel.forward_flops(counts))
        bwd = int(costmodel.backward_flops(counts))

        md_steps = 4 * args.steps
        md = {}
        for name, prec in (("fp64", engine.PRECISION_FP64),
                           ("mixed", engine.PRECISION_MIXED)):
            md[name] = run_md(model, r, elem, box, prec, md_steps,
                              args.warmup,
                              tuned=fixed.get(name, name in tuning))
            log("MD %s: %s" % (name, md[name]))
        md["mixed_vs_plain_fp64"] = (md["mixed"]["atom_steps_per_s"]
                                     / md["fp64"]["atom_steps_per_s"])
        sur_model, hist, ttrain = train_surrogate(tmp, log)
        batch = run_batch(tmp, args.steps, args.warmup, log,
                          sweep=not args.no_tune)
        md["surrogate"] = {
            "what": "radial-only HDNNP, 8 G2 per element (4 eta x 2 "
                    "neighbour elements), 8-16-16-1 softplus, trained on the "
                    "device (forward, forward-over-reverse gradient and Adam "
                    "kernels) on 64 frames of the full model's own 300 K MD "
                    "of a 1536-atom box, 20 epochs",
            "train_seconds": ttrain,
            "train_loss_first_epoch": list(hist[0]),
            "train_loss_last_epoch": list(hist[-1]),
            "multiple_time_step": "measured by bench.py --path mts (config "
                                  "C4: trained surrogate, r-RESPA with a "
                                  "force-RMSE gate)"}
        md["what"] = ("NVE velocity Verlet on the same box, dt 0.25 fs, "
                      "300 K Maxwell-Boltzmann start, 1.0 A Verlet skin "
                      "(rebuild when an atom moved > 0.5 A), list filtered "
                      "to rc every step, each step two HIP graphs")

        cpu = None
        if not args.no_cpu_baseline:
            log("cpu baseline (oracle, 24,576-atom sample)")
            cpu = cpu_baseline(tmp)
            log("cpu baseline", cpu)

        out = {
            "metric": "E+F atom-steps/s, mixed-precision vs plain FP64 GPU, "
                      "at force err <=1 meV/A",
            "value": n / dtm,
            "unit": "atom-steps/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": dtm * 1e3,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f32+f64",
            "data": "synthetic",
            "config": {"workload": "water %d atoms (%d H2O, cubic %.1f A), "
                                   "40 SFs/element (G2+G3), rc 6.35 A cos "
                                   "cutoff, nsf-25-25-1 tanh, neighbour "
                                   "list rebuilt every step; value = mixed-"
                                   "precision path (FP32 SFs and force "
                                   "contraction, FP32 MFMA networks, FP64 "
                                   "scaling, energy and force sums)"
                                   % (n, args.nmol, args.box),
                       "atoms": n, "path": "mixed",
                       "plain_fp64": {"atom_steps_per_s": n / dt64,
                                      "ms_per_step": dt64 * 1e3,
                                      "kernel_ms": kt64,
                                      "roofline": roofline(kt64, counts, nnf,
                                                           n, False, sp64)},
                       "mixed16": {"atom_steps_per_s": n / dt16,
                                   "ms_per_step": dt16 * 1e3,
                                   "kernel_ms": kt16,
                                   "error_vs_fp64": err16,
                                   "what": "as the mixed path with FP16 "
                                           "MFMA network layers (operands "
                                           "split hi + lo, three MFMAs per "
                                           "step, FP32 accumulation)"},
                       "mixed_bf16": {"atom_steps_per_s": n / dtb,
                                      "ms_per_step": dtb * 1e3,
                                      "kernel_ms": ktb,
                                      "error_vs_fp64": errbf,
                                      "what": "as the mixed path with BF16 "
                                              "MFMA network layers (operands "
                                              "split hi + lo, three MFMAs per "
                                              "step, FP32 accumulation)"},
                       "mixed_host_io": {
                           "atom_steps_per_s": n / dth,
                           "ms_per_step": dth * 1e3,
                           "forces_equal_device_path": bool(
                               np.array_equal(fh, fm)),
                           "what": "as the mixed path with the positions "
                                   "copied in from host memory and the "
                                   "forces copied out to it every step "
                                   "(PCIe included; not the value)"},
                       "mixed_vs_plain_fp64": dt64 / dtm,
                       "tuning": tuning,
                       "md": md,
                       "batch_512_structures": batch,
                       "mixed_error_vs_fp64": err,
                       "other_kernels": other_rooflines(ktm, kt64, counts,
                                                        nnf, n),
                       "flops_per_atom_step": (fwd + bwd + nnf) / n,
                       "work_counts": counts,
                       "kernel_ms": ktm},
            "roofline": roofline(ktm, counts, nnf, n, True, spm),
            "cpu_baseline": cpu,
        }
        print(json.dumps(out), flush=True)
        ctx.close()


if __name__ == "__main__":
    main()
