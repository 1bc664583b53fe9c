"""Twisted column at full size: 256 x 256 x 896 cells (dx = 1/128), 100 663 296
particles, cubic B-splines, FP64 (scenes.twisted_column(1.0)).

    python -m exagoop_amd.column [--res 1.0] [--steps 20] [--warmup 3]

prints one JSON object: particle-updates/s of the plain step over the timed
steps (HIP events on the context's stream), the per-kernel times of a
profiled pass, and size-independent checks of the state the run leaves:
the live particle count, the total mass, the net x / y momentum (zero for the
twist field by symmetry) against sqrt(2 M E_kin), finite energies.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import capi, scenes

# Algorithmic bytes per particle and per node of the USL step (FP64).
STEP_B_PER_P, STEP_B_PER_N = 640, 208
PART_B_PER_P, PART_B_PER_N = 312 + 216, 48


def run(res=1.0, steps=20, warmup=3, omega0=10.0, log=None):
    def say(msg):
        if log:
            log(msg)

    t0 = time.perf_counter()
    s, pos, r, idata = scenes.twisted_column(res, omega0=omega0)
    npart = pos.shape[1]
    nn = (s.ncells[0] + 1) * (s.ncells[1] + 1) * (s.ncells[2] + 1)
    t_gen = time.perf_counter() - t0
    say(f"[column] {list(s.ncells)} cells, {npart} particles, {nn} nodes, generated in "
        f"{t_gen:.1f} s")
    t0 = time.perf_counter()
    g = capi.Context(s)
    g.upload(pos, r, idata)
    del pos, r, idata
    t_up = time.perf_counter() - t0
    e0 = g.energies()
    g.init_internal_forces()
    g.step(warmup)
    say(f"[column] uploaded in {t_up:.1f} s, {warmup} warm-up steps done")
    g.step(steps)
    ev_ms, _ = g.last_step_timing()
    g.set_option(capi.OPT_PROFILE, 1)
    g.step(3)
    _, kms = g.last_step_timing()
    g.set_option(capi.OPT_PROFILE, 0)
    e1 = g.energies()
    n1 = g.num_particles()
    t, nsteps, dt = g.time()
    g.close()
    # momentum scale: the magnitude a uniform motion of the same kinetic energy has
    pmag = float(np.sqrt(2.0 * e1[0] * e1[5]))
    part_gbs = (PART_B_PER_P * npart + PART_B_PER_N * nn) / (kms[2] * 1e-3) / 1e9
    out = {
        "workload": "twisted column plain FP64", "cells": list(s.ncells), "particles": npart,
        "nodes": nn, "order": 3, "scheme": "USL", "omega0": omega0,
        "steps": steps, "warmup": warmup, "ms_per_step": ev_ms / steps,
        "value": npart * steps / (ev_ms * 1e-3), "unit": "particle-updates/s",
        "kernels_ms": {"p2g": kms[0], "grid": kms[1], "g2p_move_stress": kms[2], "sort": kms[3]},
        "particle_kernel_gbs": part_gbs, "particle_kernel_frac": part_gbs / 8000.0,
        "step_algorithmic_gbs": (STEP_B_PER_P * npart + STEP_B_PER_N * nn) /
                                (ev_ms / steps * 1e-3) / 1e9,
        "generate_s": t_gen, "upload_s": t_up, "time": t, "last_dt": dt,
        "checks": {
            "live_particles": n1, "particles_kept": n1 == npart,
            "mass_rel_change": abs(e1[5] - e0[5]) / e0[5],
            "net_momentum_xy_rel": [abs(e1[2]) / max(pmag, 1e-300),
                                    abs(e1[3]) / max(pmag, 1e-300)],
            "energies_finite": bool(np.all(np.isfinite(e1))),
            "kinetic_energy": [float(e0[0]), float(e1[0])],
            "strain_energy": float(e1[1]),
        },
    }
    say(f"[column] {steps} steps: {ev_ms / steps:.3f} ms/step, "
        f"{out['value'] / 1e9:.3f} G particle-updates/s")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="twisted column, plain FP64 step")
    ap.add_argument("--res", type=float, default=1.0)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    a = ap.parse_args(argv)
    out = run(a.res, a.steps, a.warmup, log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
