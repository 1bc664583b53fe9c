"""Benchmark: particle-updates/s of the plain FP64 explicit MPM step.

Workload: dam break, 128x64x64 cells (dx = 1/64), 524 288 particles
(block 0.5x1x0.5, 2x2x2 per cell), Tait EOS K = 2e6, gamma = 7, mu = 1e-3,
cubic B-splines, no-slip walls, CFL 0.3, USL, FP64, synthetic scene.
One step = the full time-loop body (dt, P2G, grid update + BCs, G2P, move,
redistribute, volume, stress).  value = particles x steps / seconds of the
timed region, inputs resident in HBM.

The deterministic object runs the same dam break with the bitwise-
reproducible P2G (MPM_OPT_DETERMINISTIC) over the same steps: milliseconds
per step, the ratio to the default step, and its per-kernel times.

The cpu_baseline leg times the CPU restatement of the reference step
(oracle/, OpenMP with atomic deposits) on the same workload: 50 steps after 5
warm-up steps on the job's host cores, and a 1-thread figure on 4 steps after
one warm-up step; the
host's core count, CPU model and sockets go into the record.

The batch object runs the batched ensemble (exagoop_amd/ensemble.py): 256
independent scenes (seeds 0..255: elastic disks or dam break, 2k-64k
particles, 16^3-48^3 cells, quadratic B-splines, FP64), one persistent kernel
launch per step replayed from a HIP graph, decomposition chosen by the
library's autotuner, 500 timed steps: particle-updates/s over all scenes and
the step's algorithmic HBM bytes per second; then the same ensemble with the
bitwise-reproducible step (MPM_OPT_DETERMINISTIC).

The column object runs the twisted column at full size (256 x 256 x 896
cells, 100 663 296 particles, cubic, exagoop_amd/column.py): 20 timed steps
after 3 warm-up steps, particle-updates/s, per-kernel times, and the
size-independent checks of its final state.

The disks object runs the elastic-disk collision (two linear-elastic disks,
64x64x4 cells, 33 008 particles, quadratic B-splines, dt = 1e-4): 200 timed
steps on the GPU and the oracle over the same steps.

The redist10 object runs the same dam break with the reference's default
redistribution interval (num_redist = 10, mpm_specs.H:89): particles are
redistributed every tenth step, so between redistributions the tile kernels
run their forms for particles outside their tile.

The musl object runs the same dam break with the MUSL scheme (the second
P2G re-projects the updated particle velocities before the particle
kernel).

The thermal object runs the same dam break with the thermal module on
(uniform temperature, adiabatic walls): milliseconds per step and the
thermal stages' share of them.

The learned object compares, on the same dam break over 200 fine steps of
dt = 2.5e-5 (a quarter of the CFL-0.3 step; the plain run at that dt is the
reference), the learned coarse step (4 fine dt per step, the committed
network exagoop_amd/data/learned_dam_break.npz), the uncorrected coarse step
and plain runs at larger fixed dt: the relative L2 particle-velocity error of
each against the reference, and the speed-up of the learned form over the
plain run at the largest dt whose error does not exceed the learned form's
(learned vs plain at equal error).  It also reports the coarse step without
the network with the PIC/FLIP blend alpha^4 (the blend of the four fine steps
a coarse step stands for, MPM_OPT_COARSE_BLEND), and the learned and that
uncorrected form's errors over four times the horizon.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-updates/sec per full MPM step; learned vs plain at equal error"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes (FP64, each persistent field read/written once per kernel)
P2G_B_PER_P, P2G_B_PER_N = 112, 56
GRID_B_PER_N = 104
PART_B_PER_P, PART_B_PER_N = 312 + 216, 48


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=50)
    ap.add_argument("--cpu-warmup", type=int, default=5)
    ap.add_argument("--cpu-1t-steps", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--scheme", default="USL", choices=["USL", "MUSL"])
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--learned-steps", type=int, default=200)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-column", action="store_true")
    ap.add_argument("--column-steps", type=int, default=20)
    ap.add_argument("--no-thermal", action="store_true")
    ap.add_argument("--no-disks", action="store_true")
    ap.add_argument("--no-det", action="store_true")
    ap.add_argument("--no-redist10", action="store_true")
    ap.add_argument("--no-musl", action="store_true")
    ap.add_argument("--thermal-steps", type=int, default=50)
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this size (HBM counter "
                         "calibration for rocprof FETCH_SIZE / WRITE_SIZE passes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    from exagoop_amd import capi, scenes

    scheme = capi.USL if args.scheme == "USL" else capi.MUSL
    s, pos, r, idata = scenes.dam_break(scheme=scheme)
    npart = pos.shape[1]
    nn = (s.ncells[0] + 1) * (s.ncells[1] + 1) * (s.ncells[2] + 1)
    print(f"[bench] dam break {npart} particles, {nn} nodes, {args.scheme}", file=sys.stderr,
          flush=True)
    g = capi.Context(s)
    g.upload(pos, r, idata)
    g.init_internal_forces()
    g.step(args.warmup)
    t0 = time.perf_counter()
    g.step(args.steps)  # returns after the stream is synchronised
    wall = time.perf_counter() - t0
    ev_ms, _ = g.last_step_timing()
    secs = ev_ms / 1e3
    value = npart * args.steps / secs
    print(f"[bench] {args.steps} steps: {ev_ms:.3f} ms (events), {wall * 1e3:.3f} ms (wall)",
          file=sys.stderr, flush=True)

    # per-kernel averages on the library's stream (separate, profiled pass)
    g.set_option(capi.OPT_PROFILE, 1)
    g.step(min(args.steps, 20))
    _, kms = g.last_step_timing()
    g.set_option(capi.OPT_PROFILE, 0)
    kern = {
        "p2g": (kms[0], P2G_B_PER_P * npart + P2G_B_PER_N * nn),
        "grid": (kms[1], GRID_B_PER_N * nn),
        "g2p_move_stress": (kms[2], PART_B_PER_P * npart + PART_B_PER_N * nn),
    }
    dom = max(kern, key=lambda k: kern[k][0])
    dms, dbytes = kern[dom]
    achieved = dbytes / (dms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "exagoop_amd", "data", "hbm_traffic_dam_break.json")
    kname = {"p2g": "k_p2g_cells", "grid": "k_grid", "g2p_move_stress": "k_particle_tile"}[dom]
    if args.scheme == "USL" and os.path.exists(tfile):
        # rocprof FETCH_SIZE / WRITE_SIZE of this kernel on this workload, scaled by a
        # plain copy of known size (profiles/dam_break_hbm_traffic.md)
        traffic = json.load(open(tfile)).get(kname, {}).get("traffic_bytes")
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "kernel": dom,
                "kernels_ms": {k: round(v[0], 5) for k, v in kern.items()},
                "kernels_gbs": {k: round(v[1] / (v[0] * 1e-3) / 1e9, 2) for k, v in kern.items()},
                "sort_ms": round(kms[3], 5)}

    det = None
    if not args.no_det:
        det = deterministic_leg(s, pos, r, idata, args.warmup, args.steps, ev_ms / args.steps)

    redist10 = None
    if not args.no_redist10:
        redist10 = variant_leg(s, pos, r, idata, args.warmup, args.steps, ev_ms / args.steps,
                               "num_redist 10", num_redist=10)

    musl = None
    if not args.no_musl and args.scheme == "USL":
        from exagoop_amd.capi import MUSL
        musl = variant_leg(s, pos, r, idata, args.warmup, args.steps, ev_ms / args.steps,
                           "MUSL", stress_update_scheme=MUSL)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_leg(s, pos, r, idata, args.cpu_warmup, args.cpu_steps, args.cpu_1t_steps)

    if args.calib_copy_mib:
        calib_copy(args.calib_copy_mib)

    learned = None
    if not args.no_learned and args.scheme == "USL":
        learned = learned_leg(args.learned_steps)

    batch = None
    if not args.no_batch:
        batch = batch_leg(args.batch_scenes, args.batch_steps)

    thermal = None
    if not args.no_thermal:
        thermal = thermal_leg(s, pos, r, idata, args.thermal_steps, ev_ms / args.steps)

    disks = None
    if not args.no_disks:
        disks = disks_leg(not args.no_cpu)

    column = None
    if not args.no_column:
        from exagoop_amd import column as column_mod
        column = column_mod.run(1.0, args.column_steps, 3,
                                log=lambda m: print(m, file=sys.stderr, flush=True))

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ev_ms / args.steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "dam break plain FP64", "cells": list(s.ncells),
                   "particles": npart, "nodes": nn, "order": 3, "scheme": args.scheme,
                   "constitutive": "Tait EOS K=2e6 gamma=7 mu=1e-3", "cfl": 0.3},
        "roofline": roofline,
        "deterministic": det,
        "redist10": redist10,
        "musl": musl,
        "cpu_baseline": cpu,
        "learned": learned,
        "batch": batch,
        "column": column,
        "thermal": thermal,
        "disks": disks,
    }
    print(json.dumps(out), flush=True)


def deterministic_leg(s, pos, r, idata, warmup, nsteps, plain_ms):
    """The same dam break with MPM_OPT_DETERMINISTIC (bitwise-reproducible
    P2G: per-cell particle order, private wavefront slabs, per-node gather of
    the tile blocks): milliseconds per step over the same steps, and its
    per-kernel times (p2g includes the gather pass)."""
    from exagoop_amd import capi
    g = capi.Context(s)
    g.set_option(capi.OPT_DETERMINISTIC, 1)
    g.upload(pos, r, idata)
    g.init_internal_forces()
    g.step(warmup)
    g.step(nsteps)
    ev_ms, _ = g.last_step_timing()
    g.set_option(capi.OPT_PROFILE, 1)
    g.step(min(nsteps, 20))
    _, kms = g.last_step_timing()
    g.close()
    ms = ev_ms / nsteps
    out = {"ms_per_step": round(ms, 5), "value": round(pos.shape[1] * 1e3 / ms, 1),
           "unit": "particle-updates/s", "vs_default_time": round(ms / plain_ms, 3),
           "kernels_ms": {"p2g_and_gather": round(kms[0], 5), "grid": round(kms[1], 5),
                          "g2p_move_stress": round(kms[2], 5)},
           "sort_ms": round(kms[3], 5)}
    print(f"[bench] deterministic P2G: {ms:.4f} ms/step ({out['vs_default_time']}x the default "
          f"step), p2g+gather {kms[0]:.4f} ms, sort {kms[3]:.4f} ms", file=sys.stderr, flush=True)
    return out


def variant_leg(s, pos, r, idata, warmup, nsteps, plain_ms, label, **changes):
    """The same dam break with the specs fields in changes: milliseconds per
    step over the same steps and the per-kernel times (the particle segment
    holds every stage between the grid pass and the particle kernel's end)."""
    import copy
    from exagoop_amd import capi
    sv = copy.deepcopy(s)
    for k, v in changes.items():
        setattr(sv, k, v)
    g = capi.Context(sv)
    g.upload(pos, r, idata)
    g.init_internal_forces()
    g.step(warmup)
    g.step(nsteps)
    ev_ms, _ = g.last_step_timing()
    g.set_option(capi.OPT_PROFILE, 1)
    g.step(min(nsteps, 20))
    _, kms = g.last_step_timing()
    g.close()
    ms = ev_ms / nsteps
    out = dict(changes)
    out.update({"ms_per_step": round(ms, 5), "value": round(pos.shape[1] * 1e3 / ms, 1),
                "unit": "particle-updates/s", "vs_default_time": round(ms / plain_ms, 3),
                "kernels_ms": {"p2g": round(kms[0], 5), "grid": round(kms[1], 5),
                               "particle_segment": round(kms[2], 5)},
                "sort_ms": round(kms[3], 5)})
    print(f"[bench] {label}: {ms:.4f} ms/step ({out['vs_default_time']}x the default step), "
          f"p2g {kms[0]:.4f} ms, particle segment {kms[2]:.4f} ms", file=sys.stderr, flush=True)
    return out


def host_cpu():
    """nproc, the CPU model and the socket count of this host (lscpu)."""
    import subprocess
    info = {"nproc": os.cpu_count(), "model": None, "sockets": None,
            "threads_allowed": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() == "Model name":
                info["model"] = v.strip()
            elif k.strip() == "Socket(s)":
                info["sockets"] = int(v.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return info


def cpu_leg(s, pos, r, idata, warmup, nsteps, nsteps_1t):
    """The oracle (the CPU restatement of the reference step, OpenMP with
    atomic deposits as the reference's CPU build) on the benchmark dam break:
    warm-up steps, then nsteps timed on the cores this job may use
    (OMP_NUM_THREADS, else the affinity mask), and a 1-thread figure over
    nsteps_1t steps."""
    import oracle
    host = host_cpu()
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(env) if env.isdigit() and int(env) > 0 else host["threads_allowed"]
    npart = pos.shape[1]

    def run(nthreads, n, w):
        o = oracle.Oracle(s, threads=nthreads)
        o.upload(pos, r, idata)
        o.init_internal_forces()
        o.step(w)
        t0 = time.perf_counter()
        o.step(n)
        return time.perf_counter() - t0

    ct = run(threads, nsteps, warmup)
    c1 = run(1, nsteps_1t, 1)
    out = {"value": round(npart * nsteps / ct, 1), "unit": "particle-updates/s",
           "cores": threads, "kind": "port",
           "sample": f"full dam break ({npart} particles): {nsteps} steps timed after "
                     f"{warmup} warm-up steps, oracle OpenMP on {threads} threads, {ct:.1f} s",
           "value_1_thread": round(npart * nsteps_1t / c1, 1),
           "sample_1_thread": f"{nsteps_1t} steps after 1 warm-up step, {c1:.1f} s",
           "host": host,
           "threads_note": "the job's core share (OMP_NUM_THREADS) of the GPU host"}
    print(f"[bench] cpu baseline {out['value']:.0f} p-upd/s on {threads} threads, "
          f"{out['value_1_thread']:.0f} on 1 thread ({host})", file=sys.stderr, flush=True)
    return out


def disks_leg(with_cpu):
    """The elastic-disk collision (two linear-elastic disks, 64x64x4 cells,
    z periodic, 8 particles per cell, quadratic B-splines, fixed dt = 1e-4,
    USL, FP64): 200 timed steps after 5 warm-up steps on the GPU, and the
    oracle over the same 200 steps on the job's cores."""
    from exagoop_amd import capi, scenes

    s, pos, r, idata = scenes.elastic_disks()
    n = pos.shape[1]
    g = capi.Context(s)
    g.upload(pos, r, idata)
    g.init_internal_forces()
    g.step(5)
    g.step(200)
    ms, _ = g.last_step_timing()
    out = {"workload": "two elastic disks colliding, 64x64x4 cells, quadratic, dt 1e-4",
           "particles": n, "steps": 200, "ms_per_step": round(ms / 200, 5),
           "value": round(n * 200 / (ms * 1e-3), 1), "unit": "particle-updates/s"}
    if with_cpu:
        import oracle
        env = os.environ.get("OMP_NUM_THREADS", "")
        threads = int(env) if env.isdigit() and int(env) > 0 else len(os.sched_getaffinity(0))
        o = oracle.Oracle(s, threads=threads)
        o.upload(pos, r, idata)
        o.init_internal_forces()
        o.step(5)
        t0 = time.perf_counter()
        o.step(200)
        ct = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(n * 200 / ct, 1), "unit": "particle-updates/s",
                               "cores": threads, "kind": "port",
                               "sample": f"200 steps after 5 warm-up steps, {ct:.1f} s"}
    print(f"[bench] elastic disks: {n} particles, {out['ms_per_step'] * 1e3:.1f} us/step, "
          f"{out['value'] / 1e9:.3f} G p-upd/s", file=sys.stderr, flush=True)
    return out


def thermal_leg(s, pos, r, idata, nsteps, plain_ms):
    """The same dam break with the thermal module on (c_p, k, T uniform and
    adiabatic walls): milliseconds per step and the thermal stages' share."""
    import numpy as np
    from exagoop_amd import capi

    n = pos.shape[1]
    th = np.zeros((capi.NTHERMAL, n))
    th[capi.TH_TEMPERATURE] = 300.0
    th[capi.TH_SPECIFIC_HEAT] = 4186.0
    th[capi.TH_CONDUCTIVITY] = 0.6
    g = capi.Context(s)
    g.upload(pos, r, idata)
    g.enable_thermal()
    g.upload_thermal(th)
    g.init_internal_forces()
    g.step(3)
    g.step(nsteps)
    ms, _ = g.last_step_timing()
    per = ms / nsteps
    print(f"[bench] thermal dam break: {per:.4f} ms/step", file=sys.stderr, flush=True)
    return {"value": round(n * nsteps / (ms * 1e-3), 1), "unit": "particle-updates/s",
            "ms_per_step": round(per, 5), "plain_ms_per_step": round(plain_ms, 5),
            "thermal_ms_per_step": round(per - plain_ms, 5), "steps": nsteps,
            "workload": "dam break with the thermal module (USE_TEMP stages), uniform T"}


def calib_copy(mib):
    """One device-to-device copy of `mib` MiB through the HIP runtime (a plain
    copy kernel of known bytes for scaling rocprof's HBM counters)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib * (1 << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) or \
            hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    hip.hipDeviceSynchronize()
    rc = hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    if rc:
        raise RuntimeError(f"hipMemcpy failed ({rc})")
    print(f"[bench] calibration copy {mib} MiB read + {mib} MiB written", file=sys.stderr,
          flush=True)


def batch_leg(nscenes, nsteps):
    """The batched ensemble: autotune, warm up (graph capture), time nsteps."""
    import numpy as np
    from exagoop_amd import capi, ensemble

    t0 = time.perf_counter()
    b, sc = ensemble.make_batch(range(nscenes))
    t_build = time.perf_counter() - t0
    best, tune_ms = b.autotune(trial_steps=3)
    b.step(10)
    t0 = time.perf_counter()
    b.step(nsteps)
    wall = time.perf_counter() - t0
    st = b.stats()
    npart = st["particles"]
    secs = st["step_ms"] * nsteps / 1e3
    # Nodes holding mass at the end of the run stand for the nodes a step touches.
    active = sum(int(np.count_nonzero(b.nodal(k, capi.N_MASS))) for k in range(nscenes))
    bytes_step = ((P2G_B_PER_P + PART_B_PER_P) * npart +
                  (P2G_B_PER_N + GRID_B_PER_N + PART_B_PER_N) * active)
    achieved = bytes_step / (st["step_ms"] * 1e-3) / 1e9
    kinds = [x[4]["kind"] for x in sc]
    out = {"value": round(npart * nsteps / secs, 1), "unit": "particle-updates/s",
           "ms_per_step": round(st["step_ms"], 5), "steps": nsteps,
           "wall_value": round(npart * nsteps / wall, 1),
           "scenes": nscenes, "particles": npart, "active_nodes": active,
           "disks": kinds.count("disks"), "dam_breaks": kinds.count("dam_break"),
           "tasks_per_step": st["tasks"], "tuning": best,
           "autotune_best_ms": round(tune_ms, 5), "build_s": round(t_build, 2),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                        "traffic": None, "kernel": "k_batch_step",
                        "algorithmic_bytes_per_step": bytes_step},
           "config": "seeds 0..%d of exagoop_amd.ensemble: quadratic, USL, CFL 0.3, FP64, one "
                     "persistent launch per step, hipGraph replay" % (nscenes - 1)}
    print(f"[bench] batch: {nscenes} scenes, {npart} particles, {st['step_ms']:.3f} ms/step, "
          f"{out['value'] / 1e9:.3f} G p-upd/s, tuning {best}", file=sys.stderr, flush=True)
    b.close()
    # the same ensemble with the bitwise-reproducible step (MPM_OPT_DETERMINISTIC)
    bd, _ = ensemble.make_batch(range(nscenes), deterministic=True)
    best_d, _ = bd.autotune(trial_steps=3)
    bd.step(10)
    bd.step(nsteps)
    std = bd.stats()
    out["deterministic"] = {"ms_per_step": round(std["step_ms"], 5),
                            "value": round(npart * 1e3 / std["step_ms"], 1),
                            "ratio": round(std["step_ms"] / st["step_ms"], 3),
                            "tuning": best_d}
    print(f"[bench] batch, deterministic step: {std['step_ms']:.3f} ms/step "
          f"({out['deterministic']['ratio']}x)", file=sys.stderr, flush=True)
    bd.close()
    return out


def long_horizon(ctx, net, nsteps):
    """Errors against the fine run over a longer horizon (no timing): the
    learned form, and the coarse step without the network with the PIC/FLIP
    blend alpha^4."""
    import numpy as np
    from exagoop_amd import capi, learned

    g, _ = ctx(learned.DT)
    g.step(nsteps)
    vf, idf = learned.velocities(g)
    g.close()
    out = {"horizon_fine_steps": nsteps}
    for name, model, blend in (("learned", net, 0), ("uncorrected_blend", None, 1)):
        g, _ = ctx(learned.DT, blend)
        fb = capi.step_learned(g, model, nsteps // learned.NFINE, learned.NFINE)
        v, ids = learned.velocities(g)
        g.close()
        e = learned.rel_l2(v, vf) if np.array_equal(ids, idf) else float("inf")
        out[name + "_rel_l2_velocity_err"] = float(f"{e:.4g}")
        out[name + "_fallbacks"] = fb
    print(f"[bench] learned over {nsteps} fine steps: error {out['learned_rel_l2_velocity_err']} "
          f"({out['learned_fallbacks']} fallbacks), uncorrected with alpha^4 blend "
          f"{out['uncorrected_blend_rel_l2_velocity_err']}", file=sys.stderr, flush=True)
    return out


def learned_leg(nsteps):
    """Learned vs plain at equal error, on the benchmark dam break.

    The reference is the plain run at the fine fixed dt = 2.5e-5 (CFL about
    0.075 at rest, a quarter of the CFL-0.3 step) over nsteps fine steps.
    The learned form takes coarse steps of 4 fine dt (1e-4, CFL about 0.3)
    with the committed network; the uncorrected form the same coarse steps
    without it.  Plain runs at fixed dt = T / n over the same horizon T give
    the plain error at each dt; the equal-error plain run is the one with the
    largest dt whose error does not exceed the learned form's, and the
    equal-error speed-up is its time over the learned form's time.  Each time
    is the median of three runs from fresh contexts, after an untimed warm-up
    of both paths."""
    import numpy as np
    from exagoop_amd import capi, learned, scenes

    def ctx(dt, blend=0):
        s, pos, r, idata = scenes.dam_break()
        s.fixed_timestep = 1
        s.timestep = dt
        g = capi.Context(s)
        g.upload(pos, r, idata)
        g.init_internal_forces()
        g.set_option(capi.OPT_COARSE_BLEND, blend)
        g.time()
        return g, pos.shape[1]

    def timed(dt, n, fn, reps=3, blend=0):
        # median wall time of reps runs from fresh contexts (each run takes a
        # host round trip per learned step, so single runs jitter)
        ts = []
        for _ in range(reps):
            g, npart = ctx(dt, blend)
            t0 = time.perf_counter()
            extra = fn(g, n)
            g.time()  # synchronises
            ts.append(time.perf_counter() - t0)
            v, ids = learned.velocities(g)
            g.close()
        return float(np.median(ts)), v, ids, npart, extra

    net = learned.load_net()
    nf = learned.NFINE
    nsteps -= nsteps % nf
    # untimed warm-up of the plain and learned paths (first launches load code)
    g, _ = ctx(learned.DT)
    g.step(nf)
    capi.step_learned(g, net, 2, nf)
    g.time()
    g.close()
    horizon = nsteps * learned.DT
    c_rest = (2e6 / 1000.0) ** 0.5 / (1.0 / 64.0)  # sound speed / dx
    t_fine, vf, idf, npart, _ = timed(learned.DT, nsteps, lambda g, n: g.step(n))

    def err(v, ids):
        return learned.rel_l2(v, vf) if np.array_equal(ids, idf) else float("inf")

    res = {}
    for name, model, blend in (("learned", net, 0), ("coarse", None, 0),
                               ("coarse_blend", None, 1)):
        t, v, ids, _, fb = timed(learned.DT, nsteps // nf,
                                 lambda g, n, m=model: capi.step_learned(g, m, n, nf),
                                 blend=blend)
        res[name] = (t, err(v, ids), fb)
    t_l, err_l, fb_l = res["learned"]
    sweep = []
    # dt = horizon / n, from just above the fine dt to the CFL-0.3 step
    for n in (190, 180, 170, 160, 128, 100, 80, 64, 50):
        dt = horizon / n
        t, v, ids, _, _ = timed(dt, n, lambda g, k: g.step(k))
        sweep.append({"dt": dt, "cfl_at_rest": round(dt * c_rest, 4), "steps": n,
                      "rel_l2_velocity_err": float(f"{err(v, ids):.4g}"),
                      "seconds": round(t, 5)})
    ok = [x for x in sweep if x["rel_l2_velocity_err"] <= err_l]
    if ok:
        eq = max(ok, key=lambda x: x["dt"])
        t_eq = eq["seconds"]
    else:  # no coarser plain run is as accurate: the fine run itself
        eq = {"dt": learned.DT, "cfl_at_rest": round(learned.DT * c_rest, 4), "steps": nsteps,
              "rel_l2_velocity_err": 0.0, "seconds": round(t_fine, 5)}
        t_eq = t_fine
    out = {"value": round(npart * nsteps / t_l, 1), "unit": "particle-updates/s",
           "speedup_equal_error": round(t_eq / t_l, 3),
           "equal_error_plain": eq,
           "plain_dt_sweep": sweep,
           "plain_value": round(npart * nsteps / t_fine, 1),
           "speedup_vs_fine": round(t_fine / t_l, 3),
           "rel_l2_velocity_err": float(f"{err_l:.4g}"),
           "uncorrected_rel_l2_velocity_err": float(f"{res['coarse'][1]:.4g}"),
           "uncorrected_value": round(npart * nsteps / res["coarse"][0], 1),
           "uncorrected_blend_rel_l2_velocity_err": float(f"{res['coarse_blend'][1]:.4g}"),
           "uncorrected_blend_note": "the coarse step without the network, PIC/FLIP blend "
                                     "alpha^4 (MPM_OPT_COARSE_BLEND)",
           "long_horizon": long_horizon(ctx, net, 4 * nsteps),
           "fallbacks": fb_l, "fine_dt": learned.DT,
           "fine_dt_note": "reference run at a quarter of the CFL-0.3 step (CFL about 0.075)",
           "fine_steps_per_learned_step": nf, "horizon_fine_steps": nsteps,
           "horizon_s": horizon,
           "network": "27x4 -> 64 -> 64 -> 3 GELU, BF16 MFMA / FP32 accumulate",
           "weights": "exagoop_amd/data/learned_dam_break.npz"}
    print(f"[bench] learned: rel L2 velocity error {out['rel_l2_velocity_err']} "
          f"(uncorrected {out['uncorrected_rel_l2_velocity_err']}, with alpha^4 blend "
          f"{out['uncorrected_blend_rel_l2_velocity_err']}); equal-error plain dt "
          f"{eq['dt']:.4g} (error {eq['rel_l2_velocity_err']}): {out['speedup_equal_error']}x; "
          f"{out['speedup_vs_fine']}x the fine run", file=sys.stderr, flush=True)
    for x in sweep:
        print(f"[bench]   plain dt {x['dt']:.4g} (CFL {x['cfl_at_rest']}): error "
              f"{x['rel_l2_velocity_err']}, {x['seconds'] * 1e3:.2f} ms", file=sys.stderr,
              flush=True)
    return out


if __name__ == "__main__":
    main()
