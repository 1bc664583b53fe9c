"""Learned coarse-step correction for the dam break: training-set generation
from the plain solver, training on the GPU, and evaluation against the plain
fine-step run.

Scenes: the dam break of scenes.dam_break at 128 x 64 x 64 cells with a
water block [0, w] x [0, 1] x [0, h], 2 x 2 x 2 particles per cell.  The
fine (plain) run uses a fixed dt = 2.5e-5, a quarter of the CFL 0.3 step at
rest; the learned coarse step covers four of them (dt = 1e-4, CFL 0.3).  A
coarse step of 4e-4 (CFL 1.2) is unstable for this fluid with or without a
nodal correction.  Training seeds 0..63 and
held-out seeds 1000..1015 draw w ~ U(0.3, 0.7) and h ~ U(0.5, 1.0) from
numpy's default_rng(seed).  One learned step covers NFINE = 4 fine steps.

Feature scales: m_ref = rho dx^3 (the mass of a full cell), v_ref = 1 m/s;
dv_ref = rms of the training targets.

    python -m exagoop_amd.learned --out DIR [--train-seeds 64] [--eval-seeds 16]

Training states are taken every 40 fine steps over the first 800 (samples =
20), so the network sees the flow as far as the long evaluation horizon; a
network trained on the first 400 steps only had held-out errors of 10-70 %
with dozens of guard fallbacks at 800 steps.

writes DIR/learned_dam_break.npz (parameters + scales) and
DIR/learned_dam_break.json (training losses; per held-out seed: relative L2
velocity error of the learned and of the uncorrected coarse run against the
fine run, fallbacks, wall times and the speedup).

The twisted column (scenes.twisted_column, full size 256 x 256 x 896 cells,
100.7 M particles, cubic) is the second problem:

    python -m exagoop_amd.learned --problem column --out DIR [--res 1.0]

Training seeds 0..31 and held-out seeds 1000..1003 draw the twist amplitude
omega0 ~ U(5, 15) rad/s from numpy's default_rng(seed).  The fine run uses a
fixed dt = 1.6e-3 dx (CFL ~ 0.075 for E = 1e6 and the fastest initial
particle); one learned step covers NFINE = 4 of them (CFL ~ 0.3).  m_ref =
rho dx^3, v_ref = 1 m/s.  Velocities are read back one field at a time by
upload index (mpm_get_particle_field), so evaluating 100 M particles costs no
full download.
"""
import argparse
import json
import os
import time

import numpy as np

from . import capi, scenes

NFINE = 4
DT = 2.5e-5
RHO = 1000.0
DX = 1.0 / 64.0
M_REF = RHO * DX ** 3
V_REF = 1.0
TRAIN_SEEDS = list(range(64))
EVAL_SEEDS = list(range(1000, 1016))
DATA_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                         "learned_dam_break.npz")


class DamBreak:
    """The dam-break problem (module constants above)."""
    name = "dam_break"

    def __init__(self):
        self.dt, self.m_ref, self.v_ref = DT, M_REF, V_REF
        self.train_seeds, self.eval_seeds = TRAIN_SEEDS, EVAL_SEEDS
        self.data_file = DATA_FILE

    def scene(self, seed):
        return scene(seed)

    def describe(self, seed):
        return {"geometry": geometry(seed)}


class Column:
    """The twisted column at resolution res (1 = full size)."""
    name = "column"

    def __init__(self, res=1.0):
        self.res = res
        self.dx = 1.0 / (128.0 * res)
        self.dt = 1.6e-3 * self.dx
        self.m_ref = RHO * self.dx ** 3
        self.v_ref = 1.0
        self.train_seeds = list(range(32))
        self.eval_seeds = list(range(1000, 1004))
        # where a column network would be read from; none ships, because the
        # trained one's held-out error was worse than the uncorrected coarse
        # step's (--out DIR writes DIR/learned_column.npz)
        self.data_file = os.path.join(os.path.dirname(DATA_FILE), "learned_column.npz")

    @staticmethod
    def omega(seed):
        return float(np.random.default_rng(seed).uniform(5.0, 15.0))

    def scene(self, seed):
        s, pos, r, idata = scenes.twisted_column(self.res, omega0=self.omega(seed))
        s.fixed_timestep = 1
        s.timestep = self.dt
        return s, pos, r, idata

    def describe(self, seed):
        return {"omega0": self.omega(seed), "res": self.res}


DAM_BREAK = DamBreak()


def problem(name, res=1.0):
    return DAM_BREAK if name == "dam_break" else Column(res)


def prob_context(prob, seed):
    """Context of problem prob's scene `seed`, uploaded and initialised; also
    the uploaded particle count."""
    s, pos, r, idata = prob.scene(seed)
    n = pos.shape[1]
    g = capi.Context(s)
    g.upload(pos, r, idata)
    del pos, r, idata
    g.init_internal_forces()
    return g, n


def velocities_by_id(g, n):
    """(3, n) particle velocities by upload index (NaN for removed particles)."""
    return np.stack([g.particle_field(capi.R_VEL + d, n) for d in range(3)])


def geometry(seed):
    rng = np.random.default_rng(seed)
    return float(rng.uniform(0.3, 0.7)), float(rng.uniform(0.5, 1.0))


def scene(seed, nx=128, ny=64, nz=64):
    w, h = geometry(seed)
    s, pos, r, idata = scenes.dam_break(nx=nx, ny=ny, nz=nz, width=w, height=h)
    s.fixed_timestep = 1
    s.timestep = DT
    return s, pos, r, idata


def context(seed, **kw):
    s, pos, r, idata = scene(seed, **kw)
    g = capi.Context(s)
    g.upload(pos, r, idata)
    g.init_internal_forces()
    return g


def collect(seeds, samples=20, spacing=40, rows=8000, capacity=6_000_000, log=print,
            prob=None, **kw):
    """Training rows from the plain trajectories of the given seeds: at
    `samples` points `spacing` fine steps apart, one collection of up to
    `rows` nodes each (the collection itself advances NFINE steps)."""
    prob = prob or DAM_BREAK
    net = capi.MLP(np.zeros(capi.MLP_NPARAMS, np.float32), scales=(prob.m_ref, prob.v_ref, 1.0))
    ds = capi.Dataset(capacity)
    for k, seed in enumerate(seeds):
        g = context(seed, **kw) if prob is DAM_BREAK else prob_context(prob, seed)[0]
        for j in range(samples):
            if spacing:
                g.step(spacing)
            capi.learn_collect(g, net, ds, NFINE, rows, seed * 1000 + j)
        g.close()
        if log and (k % 8 == 7 or k == len(seeds) - 1):
            log(f"[learned] collected seeds {k + 1}/{len(seeds)}: {ds.size()} rows")
    return ds


def train(ds, epochs=24, batch=4096, lr=2e-3, seed=0, log=print, prob=None):
    prob = prob or DAM_BREAK
    rms = ds.target_rms()
    rng = np.random.default_rng(seed)
    # He-style initialisation, zero biases
    p = np.zeros(capi.MLP_NPARAMS, np.float32)
    o = 0
    for fan_in, fan_out in ((108, 64), (64, 64), (64, 3)):
        n = fan_in * fan_out
        p[o:o + n] = (rng.standard_normal(n) * np.sqrt(1.0 / fan_in)).astype(np.float32)
        o += n + fan_out
    net = capi.MLP(p, scales=(prob.m_ref, prob.v_ref, rms))
    losses = []
    # step schedule: half the epochs at lr, 3/10 at lr / 5, the rest at lr / 25
    e1, e2 = epochs // 2, (epochs * 3) // 10
    for phase_lr, phase_ep in ((lr, e1), (lr * 0.2, e2), (lr * 0.04, epochs - e1 - e2)):
        if phase_ep <= 0:
            continue
        # in chunks of at most 20 epochs, each reported (the Adam step count
        # carries over between calls; each chunk draws its own shuffles)
        first = len(losses)
        done = 0
        while done < phase_ep:
            n = min(20, phase_ep - done)
            l = net.train(ds, n, batch, phase_lr, seed=seed + len(losses))
            losses += [float(x) for x in l]
            done += n
            if log:
                log(f"[learned] lr {phase_lr:g}: {done}/{phase_ep} epochs, loss {l[-1]:.4f}")
        if log:
            log(f"[learned] lr {phase_lr:g}: loss {losses[first]:.4f} -> {losses[-1]:.4f}")
    return net, losses


def velocities(g):
    pos, r, idata, ids = g.download()
    order = np.argsort(ids)
    return r[capi.R_VEL:capi.R_VEL + 3, order], ids[order]


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def evaluate(net, seeds, nsteps=400, guards=(0.05, 0.5), log=print, prob=None, **kw):
    """Fine plain run of nsteps vs nsteps/NFINE learned and uncorrected
    coarse steps from the same initial state; the uncorrected steps both with
    the plain PIC/FLIP blend alpha ("coarse") and with alpha^NFINE
    ("coarse_blend", MPM_OPT_COARSE_BLEND)."""
    prob = prob or DAM_BREAK
    out = []
    for seed in seeds:
        if prob is DAM_BREAK and kw:
            g, n = context(seed, **kw), None
        else:
            g, n = prob_context(prob, seed)
        capi_sync(g)
        t0 = time.perf_counter()
        g.step(nsteps)
        capi_sync(g)
        t_fine = time.perf_counter() - t0
        if n is None:
            vf, idf = velocities(g)
        else:
            vf = velocities_by_id(g, n)
        g.close()
        res = {"seed": seed, "steps": nsteps, "fine_s": t_fine, **prob.describe(seed)}
        for name, model, blend in (("learned", net, 0), ("coarse", None, 0),
                                   ("coarse_blend", None, 1)):
            if n is None:
                g = context(seed, **kw)
            else:
                g, _ = prob_context(prob, seed)
            g.set_option(capi.OPT_COARSE_BLEND, blend)
            capi_sync(g)
            t0 = time.perf_counter()
            fb = capi.step_learned(g, model, nsteps // NFINE, NFINE, guards)
            capi_sync(g)
            res[name + "_s"] = time.perf_counter() - t0
            if n is None:
                v, ids = velocities(g)
                same = ids.shape == idf.shape and np.array_equal(ids, idf)
            else:
                v = velocities_by_id(g, n)
                same = np.array_equal(np.isnan(v), np.isnan(vf))
                v, vf_ = np.nan_to_num(v), np.nan_to_num(vf)
            g.close()
            if same:
                res[name + "_err"] = rel_l2(v, vf if n is None else vf_)
            else:
                res[name + "_err"] = float("inf")
            res[name + "_fallbacks"] = fb
        res["speedup"] = res["fine_s"] / res["learned_s"]
        out.append(res)
        if log:
            log(f"[learned] seed {seed}, {nsteps} fine steps: err learned "
                f"{res['learned_err']:.3e} coarse {res['coarse_err']:.3e} (alpha^{NFINE} blend "
                f"{res['coarse_blend_err']:.3e}) fallbacks {res['learned_fallbacks']} speedup "
                f"{res['speedup']:.2f}")
    return out


def correction_floor(nsteps=200, modes=("none", "end", "two", "pvel", "pstate", "pnopos", "pvj",
                                        "gvj"), log=print):
    """Error floor of exact corrections of a coarse step, on the benchmark dam
    break: coarse steps of NFINE fine dt through the per-stage entry points,
    each given an exact correction of its own state, taken from NFINE fine
    steps run from that state in a second context.  Nodal modes: "none" (no
    correction), "end" (nodal velocity and FLIP increment corrected to the
    fine nodal velocity after NFINE steps, the network's form) and "two"
    (nodal velocity set to the mean of the NFINE fine nodal velocities, which
    moves the particles and updates F as the fine steps do, and the FLIP
    increment set so that the PIC/FLIP blend still ends at the last fine
    nodal velocity).  Particle modes, on top of "end": after the coarse step
    the particles take the fine run's velocity ("pvel"), its stress,
    deformation gradient, J, volume, density and pressure ("pstate"), or every
    field but the position ("pnopos").  Velocity and volume modes, on top of
    "end" and applied after the move, before the volume and stress update:
    each particle's fine-minus-coarse velocity and ln J, exact ("pvj") or as
    grid fields ("gvj": deposited with the particles' mass weights, P2G in a
    scratch context, and interpolated back, G2P); F is scaled by
    exp(d ln J / 3).  Returns {mode: relative L2 particle-velocity error
    against the fine run}."""
    from .capi import N_DVEL, N_VEL, R_VEL

    stress_like = (list(range(capi.R_STRESS, capi.R_STRESS + 6)) +
                   list(range(capi.R_DEFGRAD, capi.R_DEFGRAD + 9)) +
                   [capi.R_JACOBIAN, capi.R_VOLUME, capi.R_DENSITY, capi.R_PRESSURE])

    def make(dt):
        s, pos, r, idata = scenes.dam_break()
        s.fixed_timestep = 1
        s.timestep = dt
        g = capi.Context(s)
        g.upload(pos, r, idata)
        g.init_internal_forces()
        return g, s

    def by_id(state):
        pos, r, idata, ids = state
        o = np.argsort(ids)
        return pos[:, o].copy(), r[:, o].copy(), idata[:, o].copy()

    def det_f(r):
        F = r[capi.R_DEFGRAD:capi.R_DEFGRAD + 9].reshape(3, 3, -1)
        return (F[0, 0] * (F[1, 1] * F[2, 2] - F[1, 2] * F[2, 1]) -
                F[0, 1] * (F[1, 0] * F[2, 2] - F[1, 2] * F[2, 0]) +
                F[0, 2] * (F[1, 0] * F[2, 1] - F[1, 1] * F[2, 0]))

    def through_grid(spec, pos, r, idata, fields, dt):
        # fields (k, n) -> the particles' mass-weighted nodal projection,
        # interpolated back: P2G of the fields as velocities, then G2P of
        # those nodal velocities with zero increment (v' = sum N v_node)
        k = fields.shape[0]
        rc = r.copy()
        rc[R_VEL:R_VEL + 3] = 0.0
        rc[R_VEL:R_VEL + k] = fields
        C = capi.Context(spec)
        C.upload(pos, rc, idata)
        C.reset_nodal()
        C.p2g(1, 1, 0)
        nv = [C.nodal(N_VEL + d) for d in range(3)]
        for d in range(3):
            C.set_nodal(N_VEL + d, nv[d])
            C.set_nodal(N_DVEL + d, np.zeros(C.nn))
        C.g2p(1, 0, dt)
        back = by_id(C.download())[1][capi.R_VELP:capi.R_VELP + k]
        C.close()
        return back

    F, sf = make(DT)
    F.step(nsteps)
    vf = velocities(F)
    F.close()
    alpha = sf.alpha_pic_flip
    out = {}
    for mode in modes:
        A, spec = make(NFINE * DT)
        t, dtc = 0.0, NFINE * DT
        for k in range(nsteps // NFINE):
            A.reset_nodal()
            A.p2g(1, 1, 1)
            A.backup_velocity()
            A.nodal_update(dtc)
            A.apply_nodal_bcs(dtc, t)
            fine = None
            if mode != "none":
                vc = np.stack([A.nodal(N_VEL + d) for d in range(3)])
                dvc = np.stack([A.nodal(N_DVEL + d) for d in range(3)])
                # the second context takes the particles in id order, so its
                # ids are the coarse run's
                B = capi.Context(sf)
                B.upload(*by_id(A.download()))
                vs = []
                for _ in range(NFINE):
                    B.step(1)
                    vs.append(np.stack([B.nodal(N_VEL + d) for d in range(3)]))
                fine = by_id(B.download())
                B.close()
                vend, vmean = vs[-1], sum(vs) / NFINE
                if mode == "two":
                    vn = vmean
                    dvn = dvc + (vend - vc) + ((1 - alpha) / alpha) * (vend - vmean)
                else:
                    vn, dvn = vend, dvc + (vend - vc)
                for d in range(3):
                    A.set_nodal(N_VEL + d, vn[d])
                    A.set_nodal(N_DVEL + d, dvn[d])
            A.g2p(1, 1, dtc)
            A.move(dtc, k + 1)
            if mode in ("pvj", "gvj"):
                pa, ra, ia = by_id(A.download())
                rb = fine[1]
                dv = rb[R_VEL:R_VEL + 3] - ra[R_VEL:R_VEL + 3]
                dl = (np.log(rb[capi.R_JACOBIAN]) - np.log(det_f(ra)))[None, :]
                if mode == "gvj":
                    dv = through_grid(spec, pa, ra, ia, dv, dtc)
                    dl = through_grid(spec, pa, ra, ia, dl, dtc)
                ra[R_VEL:R_VEL + 3] += dv
                ra[capi.R_DEFGRAD:capi.R_DEFGRAD + 9] *= np.exp(dl[0] / 3.0)
                A.upload(pa, ra, ia)
            A.update_volume()
            A.stress(dtc, t)
            if mode in ("pvel", "pstate", "pnopos"):
                pa, ra, ia = by_id(A.download())
                rb = fine[1]
                if mode == "pnopos":
                    ra = rb
                else:
                    for f in (range(R_VEL, R_VEL + 3) if mode == "pvel" else stress_like):
                        ra[f] = rb[f]
                A.upload(pa, ra, ia)
            t += dtc
        v = velocities(A)
        A.close()
        out[mode] = rel_l2(v[0], vf[0]) if np.array_equal(v[1], vf[1]) else float("inf")
        if log:
            log(f"[learned] exact-correction floor, {nsteps} fine steps, {mode}: "
                f"{out[mode]:.3e}")
    return out


def capi_sync(g):
    g.time()  # synchronises the context's stream


def save(net, path):
    p, sc = net.get_params()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=p, scales=sc)


def load_net(path=DATA_FILE):
    d = np.load(path, allow_pickle=False)
    return capi.MLP(d["params"], scales=d["scales"])


def demo():
    """A short run of every learned-correction kernel for profiling: 20
    learned steps of the committed network on held-out seed 1000, then row
    collection on two training seeds and three epochs of training."""
    net = load_net()
    g = context(1000)
    fb = capi.step_learned(g, net, 20, NFINE)
    g.close()
    print(f"[learned demo] 20 learned steps, {fb} fallbacks", flush=True)
    ds = collect([1, 2], samples=2, spacing=10, rows=8000, capacity=100_000)
    _, losses = train(ds, epochs=3, batch=4096)
    print(f"[learned demo] training losses {losses}", flush=True)


def eval_summary(net, seeds, steps, steps_long=0, prob=None, log=print):
    """Held-out evaluation at one or two horizons, with per-horizon means."""
    out = {}
    for key, n in (("eval", steps), ("eval_long", steps_long)):
        if not n:
            continue
        ev = evaluate(net, seeds, nsteps=n, prob=prob, log=log)
        sfx = "" if key == "eval" else "_long"
        out[key] = ev
        out["steps" + sfx] = n
        for f in ("learned_err", "coarse_err", "coarse_blend_err", "speedup"):
            out["mean_" + f + sfx] = float(np.mean([e[f] for e in ev]))
        out["fallbacks" + sfx] = int(sum(e["learned_fallbacks"] for e in ev))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--demo", action="store_true",
                    help="short run of every kernel (for profiling), no outputs")
    ap.add_argument("--out")
    ap.add_argument("--problem", default="dam_break", choices=["dam_break", "column"])
    ap.add_argument("--res", type=float, default=1.0, help="column resolution (1 = full size)")
    ap.add_argument("--train-seeds", type=int, default=None)
    ap.add_argument("--eval-seeds", type=int, default=None)
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--eval-steps", type=int, default=400)
    ap.add_argument("--eval-steps-long", type=int, default=800,
                    help="second evaluation horizon (0: none)")
    ap.add_argument("--samples", type=int, default=20)
    ap.add_argument("--spacing", type=int, default=40)
    ap.add_argument("--rows", type=int, default=16000)
    ap.add_argument("--capacity", type=int, default=24_000_000,
                    help="training rows held on the device (rows beyond it are dropped)")
    ap.add_argument("--floor", type=int, metavar="NSTEPS",
                    help="error floor of exact nodal and particle corrections over NSTEPS "
                         "fine steps (benchmark dam break), no outputs")
    ap.add_argument("--eval-only", metavar="NPZ",
                    help="evaluate this network (no collection or training)")
    a = ap.parse_args(argv)
    if a.demo:
        demo()
        return
    if a.floor:
        print(json.dumps(correction_floor(a.floor)), flush=True)
        return
    if not a.out:
        ap.error("--out is required")
    os.makedirs(a.out, exist_ok=True)
    prob = problem(a.problem, a.res)
    train_seeds = prob.train_seeds[:a.train_seeds] if a.train_seeds else prob.train_seeds
    eval_seeds = prob.eval_seeds[:a.eval_seeds] if a.eval_seeds else prob.eval_seeds
    if a.eval_only:
        summary = {"problem": prob.name, "network": os.path.basename(a.eval_only)}
        summary.update(eval_summary(load_net(a.eval_only), eval_seeds, a.eval_steps,
                                    a.eval_steps_long, prob))
        with open(os.path.join(a.out, "learned_%s_eval.json" % prob.name), "w") as f:
            json.dump(summary, f, indent=1)
        print(json.dumps({k: v for k, v in summary.items() if k not in ("eval", "eval_long")}),
              flush=True)
        return
    t0 = time.perf_counter()
    ds = collect(train_seeds, samples=a.samples, spacing=a.spacing, rows=a.rows, prob=prob,
                 capacity=a.capacity)
    t1 = time.perf_counter()
    net, losses = train(ds, epochs=a.epochs, prob=prob)
    t2 = time.perf_counter()
    save(net, os.path.join(a.out, "learned_%s.npz" % prob.name))
    summary = {
        "problem": prob.name, "train_seeds": len(train_seeds), "fine_dt": prob.dt,
        "samples": a.samples, "spacing": a.spacing,
        "rows": ds.size(), "dv_ref": float(net.get_params()[1][2]),
        "collect_s": t1 - t0, "train_s": t2 - t1, "losses": losses,
    }
    summary.update(eval_summary(net, eval_seeds, a.eval_steps, a.eval_steps_long, prob))
    with open(os.path.join(a.out, "learned_%s.json" % prob.name), "w") as f:
        json.dump(summary, f, indent=1)
    print(json.dumps({k: v for k, v in summary.items()
                      if k not in ("eval", "eval_long", "losses")}), flush=True)


if __name__ == "__main__":
    main()
