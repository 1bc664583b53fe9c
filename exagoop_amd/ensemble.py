"""Batched ensembles of small independent scenes (the mpm_batch_* entry
points of libexagoop_amd).

Scene s of an ensemble is drawn from numpy's default_rng(seed0 + s):
  kind       elastic disks or dam break, one each with probability 1/2
  particles  target count log-uniform in [2048, 65536]
  grid       n x n x n cells on the unit cube, n uniform in {16, ..., 48}
  material   E log-uniform in [1e3, 1e5] (the disks' Young's modulus, the
             fluid's bulk modulus K), nu uniform in [0.2, 0.4], rho = 1000
  motion     initial speed uniform in [0, 0.5] along a uniformly random
             direction (the two disks move in opposite directions)
Common: quadratic B-splines, USL, CFL 0.3 (the dt follows each scene's
stiffness and speed), redistribution every step.

  disks      two linear-elastic balls of equal radius centred at
             (0.3, 0.3, 0.5) and (0.7, 0.7, 0.5), slip walls, no gravity
  dam break  Tait fluid (K = E, gamma = 7, mu = 1e-3) in the block
             [0, a] x [0, 1] x [0, h] against the x-low wall, gravity -9.81
             along z, no-slip walls
The shape's volume fraction is set so that a lattice of p x p x p particles
per cell (p >= 2, the smallest that fits the target with the shape inside the
box) holds about the target count; the count used is the lattice's.
"""
import math

import numpy as np

from . import capi, scenes
from .capi import BC_NOSLIP, BC_SLIP, USL, default_specs

MIN_PARTICLES, MAX_PARTICLES = 2048, 65536
MIN_CELLS, MAX_CELLS = 16, 48
BALL_CENTRES = ((0.3, 0.3, 0.5), (0.7, 0.7, 0.5))
MAX_BALL_FRACTION = 2 * (4.0 / 3.0) * math.pi * 0.25 ** 3  # two balls of radius 0.25
MAX_BLOCK_FRACTION = 0.45


def draw(seed):
    """The random parameters of scene `seed`."""
    rng = np.random.default_rng(seed)
    kind = "disks" if rng.integers(2) == 0 else "dam_break"
    target = int(round(math.exp(rng.uniform(math.log(MIN_PARTICLES), math.log(MAX_PARTICLES)))))
    n = int(rng.integers(MIN_CELLS, MAX_CELLS + 1))
    E = float(math.exp(rng.uniform(math.log(1e3), math.log(1e5))))
    nu = float(rng.uniform(0.2, 0.4))
    speed = float(rng.uniform(0.0, 0.5))
    d = rng.standard_normal(3)
    d /= np.linalg.norm(d)
    return {"kind": kind, "target": target, "n": n, "E": E, "nu": nu, "speed": speed,
            "direction": d.tolist()}


def _lattice_box(n, p, lo, hi, inside):
    """Particles of the p^3-per-cell lattice of the unit cube's n^3 cells that
    lie in the cells overlapping [lo, hi] and satisfy inside()."""
    dx = 1.0 / n
    offs = (2 * np.arange(1, p + 1) - 1) / (2 * p)
    coords = []
    for d in range(3):
        c0 = max(0, int(math.floor(lo[d] / dx)))
        c1 = min(n, int(math.ceil(hi[d] / dx)))
        c = (np.arange(c0, c1)[:, None] + offs[None, :]).ravel() * dx
        coords.append(c)
    X, Y, Z = np.meshgrid(coords[0], coords[1], coords[2], indexing="ij")
    X, Y, Z = X.ravel(), Y.ravel(), Z.ravel()
    keep = inside(X, Y, Z)
    return np.stack([X[keep], Y[keep], Z[keep]]), dx ** 3 / p ** 3


def scene(seed):
    """(specs, pos, rdata, idata, params) of scene `seed`."""
    prm = draw(seed)
    n, target = prm["n"], prm["target"]
    fmax = MAX_BALL_FRACTION if prm["kind"] == "disks" else MAX_BLOCK_FRACTION
    p = 2
    while target > fmax * (p * n) ** 3:
        p += 1
    frac = target / float((p * n) ** 3)
    s = default_specs()
    s.plo[:] = [0.0, 0.0, 0.0]
    s.phi[:] = [1.0, 1.0, 1.0]
    s.ncells[:] = [n, n, n]
    s.periodic[:] = [0, 0, 0]
    s.order_scheme = 2
    s.num_redist = 1
    s.fixed_timestep = 0
    s.CFL = 0.3
    s.stress_update_scheme = USL
    vdir = np.asarray(prm["direction"]) * prm["speed"]
    if prm["kind"] == "disks":
        s.gravity[:] = [0.0, 0.0, 0.0]
        for d in range(3):
            s.bclo[d] = s.bchi[d] = BC_SLIP
        r = (3.0 * frac / (8.0 * math.pi)) ** (1.0 / 3.0)
        parts = []
        for c in BALL_CENTRES:
            lo = [c[d] - r for d in range(3)]
            hi = [c[d] + r for d in range(3)]

            def inside(X, Y, Z, c=c):
                return (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2 <= r * r

            parts.append(_lattice_box(n, p, lo, hi, inside))
        pos = np.concatenate([parts[0][0], parts[1][0]], axis=1)
        vp = parts[0][1]
        vel = np.zeros_like(pos)
        n0 = parts[0][0].shape[1]
        vel[:, :n0] = vdir[:, None]
        vel[:, n0:] = -vdir[:, None]
        rdata, idata = scenes.make_records(pos, vp, 1000.0, 0, E=prm["E"], nu=prm["nu"], vel=vel)
        prm["radius"] = r
    else:
        s.gravity[:] = [0.0, 0.0, -9.81]
        for d in range(3):
            s.bclo[d] = s.bchi[d] = BC_NOSLIP
        h = min(0.9, math.sqrt(2.0 * frac))
        a = frac / h

        def inside(X, Y, Z):
            return (X <= a) & (Z <= h)

        pos, vp = _lattice_box(n, p, [0.0, 0.0, 0.0], [a, 1.0, h], inside)
        vel = np.repeat(vdir[:, None], pos.shape[1], axis=1)
        rdata, idata = scenes.make_records(pos, vp, 1000.0, 1, K=prm["E"], gamma=7.0, mu=1e-3,
                                           vel=vel)
        prm["block"] = [a, 1.0, h]
    prm["ppc"] = p
    prm["particles"] = int(pos.shape[1])
    return s, pos, rdata, idata, prm


def make_batch(seeds, tuning=None, deterministic=False):
    """A capi.Batch of the given scene seeds, uploaded and initialised
    (deterministic: MPM_OPT_DETERMINISTIC, the bitwise-reproducible step)."""
    sc = [scene(int(k)) for k in seeds]
    b = capi.Batch([x[0] for x in sc], [x[1].shape[1] for x in sc])
    for k, (s, pos, rdata, idata, _) in enumerate(sc):
        b.upload(k, pos, rdata, idata)
    if deterministic:
        b.set_option(capi.OPT_DETERMINISTIC, 1)
    if tuning:
        b.set_tuning(**tuning)
    b.init()
    return b, sc


def main(argv=None):
    """Run an ensemble: build, autotune, time, and optionally profile the
    persistent step's stages; prints one JSON object."""
    import argparse
    import json
    import time

    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--scenes", type=int, default=256)
    ap.add_argument("--seed0", type=int, default=0)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="then run this many instrumented steps and report per-stage times")
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    b, sc = make_batch(range(a.seed0, a.seed0 + a.scenes))
    out = {"scenes": a.scenes, "build_s": round(time.perf_counter() - t0, 2)}
    if not a.no_autotune:
        best, ms = b.autotune(trial_steps=3)
        out["autotune"] = {"tuning": best, "ms_per_step": ms}
    b.step(10)
    b.step(a.steps)
    st = b.stats()
    out.update({"steps": a.steps, "ms_per_step": st["step_ms"], "particles": st["particles"],
                "tasks_per_step": st["tasks"], "tuning": b.tuning(),
                "particle_updates_per_s": st["particles"] / (st["step_ms"] * 1e-3)})
    if a.profile_steps:
        out["profile"] = b.profile(a.profile_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
