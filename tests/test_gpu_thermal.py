"""GPU parity of the thermal module (the USE_TEMP build's heat conduction)
with the CPU oracle on identical inputs (FP64).

Bars: thermal particle fields and the nodal thermal components (20..25)
relative L2 <= 1e-12 after the initial pass and one step, <= 1e-10 after 20
steps (the same bars as the mechanical record; the only expected difference
is the summation order of the FP64 atomic deposits).  Cases: linear,
quadratic and cubic shape functions; USL and MUSL; the sorted and the
unsorted step; every wall condition, mixed on one box (the corner nodes of
mixed walls exercise the serial sweep order); periodic directions (the
SumBoundary of the periodic face nodes).
"""
import sys

import numpy as np
import pytest

import oracle
from exagoop_amd import capi, scenes
from exagoop_amd.capi import MUSL, USL

pytestmark = pytest.mark.gpu

TH_COMPS = (capi.N_MASS_SPHEAT, capi.N_MASS_SPHEAT_TEMP, capi.N_TEMPERATURE,
            capi.N_SOURCE_TEMP, capi.N_DELTA_TEMPERATURE, capi.N_MASS_CONDUCTIVITY)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a, dtype=float).ravel()
    b = np.asarray(b, dtype=float).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a) if nb == 0.0 else np.linalg.norm(a - b) / nb


def thermal_fields(n, seed):
    rng = np.random.default_rng(seed)
    th = np.zeros((capi.NTHERMAL, n))
    th[capi.TH_TEMPERATURE] = rng.uniform(0.0, 1.0, n)
    th[capi.TH_SPECIFIC_HEAT] = rng.uniform(1.0, 2.0, n)
    th[capi.TH_CONDUCTIVITY] = rng.uniform(20.0, 60.0, n)
    th[capi.TH_HEAT_SOURCE] = rng.uniform(-100.0, 100.0, n)
    return th


def mixed_walls(s):
    """Every temperature condition on some face of the box."""
    t = capi.default_thermal_specs(s)
    lo = [capi.BC_TEMP_ISOTHERMAL, capi.BC_TEMP_USERFLUX, capi.BC_TEMP_ADIABATIC]
    hi = [capi.BC_TEMP_CONVECTION, capi.BC_TEMP_ADIABATIC, capi.BC_TEMP_CONVECTION]
    for d in range(3):
        if not s.periodic[d]:
            t.bclo[d], t.bchi[d] = lo[d], hi[d]
        t.T_wall_lo[d], t.T_wall_hi[d] = 2.0, -0.5
        t.flux_lo[d], t.flux_hi[d] = 10.0, -4.0
        t.h_lo[d], t.h_hi[d] = 30.0, 50.0
        t.T_inf_lo[d], t.T_inf_hi[d] = -1.0, 1.5
    return t


def pair(s, pos, r, idata, th, tspecs, sorted_=1):
    g = capi.Context(s)
    g.set_option(capi.OPT_SORTED, sorted_)
    o = oracle.Oracle(s)
    for c in (g, o):
        c.upload(pos, r, idata)
        c.enable_thermal(tspecs)
        c.upload_thermal(th)
        c.init_internal_forces()
    return g, o


def compare(g, o, tol, nodal=True):
    tg, to = g.download_thermal(), o.download_thermal()
    assert np.array_equal(np.isnan(tg), np.isnan(to))
    live = ~np.isnan(to[0])
    bad = {f: rel(tg[f][live], to[f][live]) for f in range(capi.NTHERMAL)
           if not rel(tg[f][live], to[f][live]) <= tol}
    assert not bad, bad
    if nodal:
        errs = {c: rel(g.nodal(c), o.nodal(c)) for c in TH_COMPS}
        bad = {c: e for c, e in errs.items() if not e <= tol}
        assert not bad, bad


CASES = [
    # (order, scheme, sorted, periodic)
    (1, USL, 1, (0, 0, 0)),
    (2, USL, 1, (0, 0, 0)),
    (3, USL, 1, (0, 0, 0)),
    (2, USL, 0, (0, 0, 0)),
    (2, MUSL, 1, (0, 0, 0)),
    (3, MUSL, 0, (0, 0, 0)),
    (2, USL, 1, (1, 1, 0)),
    (1, MUSL, 1, (1, 0, 1)),
]


@pytest.mark.parametrize("order,scheme,sorted_,periodic", CASES)
def test_thermal_step_matches_the_oracle(gpu, order, scheme, sorted_, periodic):
    say(f"[thermal] order={order} scheme={scheme} sorted={sorted_} periodic={periodic}")
    s, pos, r, idata = scenes.random_block(71 + order, n=(12, 10, 8), order=order,
                                           periodic=periodic, model=0, scheme=scheme,
                                           frac=0.9, speed=0.2)
    th = thermal_fields(pos.shape[1], 5 + order)
    g, o = pair(s, pos, r, idata, th, mixed_walls(s), sorted_)
    compare(g, o, 1e-12)
    for c in (g, o):
        c.step(1)
    compare(g, o, 1e-12)
    for c in (g, o):
        c.step(19)
    compare(g, o, 1e-10)
    pg, _, _, _ = g.download()
    po, _, _, _ = o.download()
    assert rel(pg, po) <= 1e-10


@pytest.mark.parametrize("scheme", [USL, MUSL])
def test_face_functions_match_the_oracle(gpu, scheme):
    """Temperature functions on five faces (isothermal, user flux, two
    convection faces and an adiabatic one, where a function changes
    nothing), varying in space and time, after the wall conditions of every
    BC pass of the step."""
    say(f"[thermal] face functions scheme={scheme}")
    s, pos, r, idata = scenes.random_block(101, n=(10, 9, 8), order=2, model=0, scheme=scheme,
                                           frac=1.0, speed=0.1)
    th = thermal_fields(pos.shape[1], 17)
    g, o = pair(s, pos, r, idata, th, mixed_walls(s))
    fns = {
        (0, 0): lambda x, y, z, t: (1.0 + y + 2.0 * z + 50.0 * t, 0.0),
        (1, 0): lambda x, y, z, t: (5.0 * np.sin(8.0 * x) + t, 0.0),
        (0, 1): lambda x, y, z, t: (20.0 + 10.0 * z, y - t),
        (2, 1): lambda x, y, z, t: (15.0 + 30.0 * x * y, 2.0),
        (2, 0): lambda x, y, z, t: (99.0, 99.0),
    }
    for c in (g, o):
        for (d, hi), fn in fns.items():
            c.set_wall_temp_udf(d, hi, fn)
        c.step(1)
    compare(g, o, 1e-12)
    for c in (g, o):
        c.step(9)
    compare(g, o, 1e-10)


def trilinear(phi, s, x):
    """Nodal phi at points x (3, n), trilinear on the nodes."""
    n = np.array(s.ncells)
    lo = np.array(s.plo)
    dx = (np.array(s.phi) - lo) / n
    f = (x - lo[:, None]) / dx[:, None]
    i = np.clip(np.floor(f).astype(int), 0, n[:, None] - 1)
    t = f - i
    out = np.zeros(x.shape[1])
    for cz in (0, 1):
        for cy in (0, 1):
            for cx in (0, 1):
                w = ((t[0] if cx else 1 - t[0]) * (t[1] if cy else 1 - t[1]) *
                     (t[2] if cz else 1 - t[2]))
                out += w * phi[i[2] + cz, i[1] + cy, i[0] + cx]
    return out


@pytest.mark.parametrize("scheme", [USL, MUSL])
def test_levelset_body_conditions_match_the_oracle(gpu, scheme):
    """Three bodies with an isothermal, a heat-flux and a convective
    condition (particles inside them removed first, as
    removeParticlesInsideEB does): the body sweeps, whose nodes read
    neighbours updated earlier in the same sweep, match the oracle's serial
    sweeps."""
    say(f"[thermal] level-set body conditions scheme={scheme}")
    s, pos, r, idata = scenes.random_block(107, n=(14, 12, 10), order=2, model=0, scheme=scheme,
                                           frac=1.0, speed=0.1)
    bodies = [
        (capi.LS_SPHERE, [0.3, 0.35, 0.3, 0.1], capi.LS_TEMP_ISOTHERMAL, dict(T_wall=2.0)),
        (capi.LS_BOX, [0.6, 0.2, 0.1, 0.75, 0.4, 0.3], capi.LS_TEMP_HEATFLUX,
         dict(heat_flux=15.0)),
        (capi.LS_CYLINDER, [0.5, 0.0, 0.45, 0.08, 0.0, 1], capi.LS_TEMP_CONVECTIVE,
         dict(h_conv=60.0, T_inf=-1.0)),
    ]
    phis = [capi.levelset_analytic(s, kind, prm) for kind, prm, _, _ in bodies]
    keep = np.ones(pos.shape[1], bool)
    for phi in phis:
        keep &= trilinear(phi, s, pos) >= 1e-20
    pos, r, idata = pos[:, keep], r[:, keep], idata[:, keep]
    th = thermal_fields(pos.shape[1], 19)
    g = capi.Context(s)
    o = oracle.Oracle(s)
    for c in (g, o):
        c.upload(pos, r, idata)
        for b, (phi, (_, _, tbc, kw)) in enumerate(zip(phis, bodies)):
            c.add_levelset_body(phi, capi.BC_SLIP, 0.0)
            c.set_levelset_body_temp(b, tbc, **kw)
        c.enable_thermal(mixed_walls(s))
        c.upload_thermal(th)
        c.init_internal_forces()
        c.step(1)
    compare(g, o, 1e-12)
    for c in (g, o):
        c.step(9)
    compare(g, o, 1e-10)


# Redistribution every five steps with particles drifting through the
# periodic x faces: between redistributions particles lie outside their tile
# (and outside the domain), so the tiled thermal deposit sends them through
# its per-particle fallback.  Mixed orders: z with 4 cells makes cubic and
# quadratic fall back to linear there (per-direction orders in the tile
# deposit and G2P).
REDIST = [
    ("drift_usl_o2", dict(order=2, periodic=(1, 0, 0), speed=0.5), 3.0),
    ("drift_musl_o3", dict(order=3, periodic=(1, 0, 1), speed=0.5, scheme=MUSL), 3.0),
    ("mixed_order_usl", dict(order=3, n=(12, 10, 4), speed=0.3), 0.0),
]


@pytest.mark.parametrize("name,kw,drift", REDIST, ids=[c[0] for c in REDIST])
def test_thermal_redistribution_interval_matches_the_oracle(gpu, name, kw, drift):
    say(f"[thermal] num_redist=5 {name}")
    s, pos, r, idata = scenes.random_block(57, frac=0.9, **kw)
    s.num_redist = 5
    r[capi.R_VEL] += drift
    th = thermal_fields(pos.shape[1], 21)
    g, o = pair(s, pos, r, idata, th, mixed_walls(s))
    for c in (g, o):
        c.step(1)
    compare(g, o, 1e-12)
    outside = 0
    for _ in range(4):
        for c in (g, o):
            c.step(3)
        pg, _, _, _ = g.download()
        outside += int(np.sum((pg < np.array(s.plo)[:, None]) |
                              (pg > np.array(s.phi)[:, None])))
        for c in (g, o):
            c.step(2)
    compare(g, o, 1e-10)
    if drift:
        assert outside > 0  # particles outside their tile deposited through the fallback


@pytest.mark.parametrize("scheme", [USL, MUSL])
def test_deterministic_thermal_step(gpu, scheme):
    """MPM_OPT_DETERMINISTIC with the thermal module: the thermal deposit's
    tile blocks and gather pass in a fixed order.  Parity with the oracle
    (1e-12 after one step, 1e-10 after 20) and two runs byte-identical in the
    thermal record and the thermal nodal components."""
    say(f"[thermal] deterministic, scheme {scheme}")
    s, pos, r, idata = scenes.random_block(61, n=(12, 10, 8), order=3, periodic=(1, 0, 0),
                                           model=0, scheme=scheme, frac=0.9, speed=0.2)
    th = thermal_fields(pos.shape[1], 17)

    def det_run(nsteps):
        g = capi.Context(s)
        g.set_option(capi.OPT_DETERMINISTIC, 1)
        g.upload(pos, r, idata)
        g.enable_thermal(mixed_walls(s))
        g.upload_thermal(th)
        g.init_internal_forces()
        g.step(nsteps)
        out = (g.download_thermal(), np.stack([g.nodal(c) for c in TH_COMPS]))
        return g, out

    g, _ = det_run(0)
    o = oracle.Oracle(s)
    o.upload(pos, r, idata)
    o.enable_thermal(mixed_walls(s))
    o.upload_thermal(th)
    o.init_internal_forces()
    for c in (g, o):
        c.step(1)
    compare(g, o, 1e-12)
    for c in (g, o):
        c.step(19)
    compare(g, o, 1e-10)
    g.close()
    _, a = det_run(20)
    _, b = det_run(20)
    for x, y in zip(a, b):
        assert x.tobytes() == y.tobytes()


def test_thermal_fields_follow_the_particles_through_removals(gpu):
    """Particles leaving a non-periodic box are removed (Redistribute); the
    thermal record follows the survivors by upload index on both sides."""
    say("[thermal] removals")
    s, pos, r, idata = scenes.random_block(83, order=2, frac=0.9, speed=0.1,
                                           bcs=[(0, 0), (capi.BC_SLIP, capi.BC_SLIP),
                                                (capi.BC_SLIP, capi.BC_SLIP)])
    r[capi.R_VEL] += 3.0  # the block drifts out through the x-hi face (periodic-coded wall)
    th = thermal_fields(pos.shape[1], 9)
    g, o = pair(s, pos, r, idata, th, capi.default_thermal_specs(s))
    for c in (g, o):
        c.step(25)
    assert g.num_particles() == o.num_particles() < pos.shape[1]
    compare(g, o, 1e-10, nodal=False)


def test_corner_nodes_follow_the_serial_sweep(gpu):
    """A box filled to its walls with a different condition on each face:
    the edge and corner nodes take the values of the reference's in-place
    serial sweep (a node reads the updated value of a wall neighbour that
    comes earlier), after one step and after ten."""
    say("[thermal] corner sweep order")
    s, pos, r, idata = scenes.random_block(89, n=(6, 5, 7), order=3, model=0, frac=1.0,
                                           speed=0.05)
    t = capi.default_thermal_specs(s)
    t.bclo[:] = [capi.BC_TEMP_CONVECTION, capi.BC_TEMP_ADIABATIC, capi.BC_TEMP_USERFLUX]
    t.bchi[:] = [capi.BC_TEMP_USERFLUX, capi.BC_TEMP_CONVECTION, capi.BC_TEMP_ADIABATIC]
    for d in range(3):
        t.flux_lo[d], t.flux_hi[d] = 5.0 + d, -3.0 - d
        t.h_lo[d], t.h_hi[d] = 40.0, 25.0 + d
        t.T_inf_lo[d], t.T_inf_hi[d] = -2.0, 3.0
    th = thermal_fields(pos.shape[1], 13)
    g, o = pair(s, pos, r, idata, th, t)
    for c in (g, o):
        c.step(1)
    compare(g, o, 1e-12)
    T = o.nodal(capi.N_TEMPERATURE).reshape(8, 6, 7)  # (k, j, i)
    assert np.abs(T[0, 0, 0]) > 0 and np.abs(T[-1, -1, -1]) > 0
    for c in (g, o):
        c.step(9)
    compare(g, o, 1e-10)


def test_thermal_off_leaves_the_step_unchanged(gpu):
    """Enabling the module changes no mechanical field."""
    say("[thermal] mechanical record unchanged")
    s, pos, r, idata = scenes.random_block(97, order=2, model=0)
    a = capi.Context(s)
    b = capi.Context(s)
    for c in (a, b):
        c.upload(pos, r, idata)
    b.enable_thermal()
    b.upload_thermal(thermal_fields(pos.shape[1], 3))
    for c in (a, b):
        c.init_internal_forces()
        c.step(5)
    pa, ra, _, _ = a.download()
    pb, rb, _, _ = b.download()
    assert rel(pb, pa) <= 1e-13 and rel(rb, ra) <= 1e-12
