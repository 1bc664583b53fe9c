"""Parity of the HIP library with the CPU oracle on identical inputs (FP64).

Bars: nodal mass/momentum/force/delta-velocity and particle state relative
L2 <= 1e-12 after one step; particle positions and stresses <= 1e-10 after
200 steps; total mass bit-conserved.  The only expected difference is the
summation order of the FP64 atomic deposits.
"""
import sys

import numpy as np
import pytest

import oracle
from exagoop_amd import capi, scenes
from exagoop_amd.capi import (BC_NOSLIP, BC_PARTIALSLIP, BC_SLIP, MUSL, N_DVEL, N_FRC, N_MASS,
                              N_MASS_OLD, N_VEL, R_MASS, R_STRAINRATE, R_STRESS, R_VEL, R_VELP,
                              USL)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a, dtype=float).ravel()
    b = np.asarray(b, dtype=float).ravel()
    nb = np.linalg.norm(b)
    if nb == 0.0:
        return np.linalg.norm(a)
    return np.linalg.norm(a - b) / nb


def pair(s, pos, r, idata):
    g = capi.Context(s)
    o = oracle.Oracle(s)
    g.upload(pos, r, idata)
    o.upload(pos, r, idata)
    return g, o


def compare_nodal(g, o, tol, comps=(N_MASS, N_VEL, N_VEL + 1, N_VEL + 2, N_FRC, N_FRC + 1,
                                    N_FRC + 2, N_DVEL, N_DVEL + 1, N_DVEL + 2, N_MASS_OLD)):
    errs = {c: rel(g.nodal(c), o.nodal(c)) for c in comps}
    bad = {c: e for c, e in errs.items() if not e <= tol}
    assert not bad, bad
    return errs


def compare_particles(g, o, tol, fields=None):
    pg, rg, ig, idg = g.download()
    po, ro, io, ido = o.download()
    assert np.array_equal(idg, ido)
    assert np.array_equal(ig, io)
    assert rel(pg, po) <= tol, rel(pg, po)
    fields = range(rg.shape[0]) if fields is None else fields
    bad = {f: rel(rg[f], ro[f]) for f in fields if not rel(rg[f], ro[f]) <= tol}
    assert not bad, bad


CASES = [
    # name, scene kwargs, spec tweaks
    ("order1", dict(order=1), {}),
    ("order2", dict(order=2), {}),
    ("order3", dict(order=3), {}),
    ("order2_periodic_xz", dict(order=2, periodic=(1, 0, 1)), {}),
    ("order3_periodic_all", dict(order=3, periodic=(1, 1, 1), n=(8, 8, 8)), {}),
    ("order1_periodic_y", dict(order=1, periodic=(0, 1, 0)), {}),
    ("mixed_order_h5", dict(order=3, n=(12, 10, 4)), {}),
    ("fluid_order3", dict(order=3, model=1), {}),
    ("neohookean_order2", dict(order=2, model=2), {}),
    ("musl_order2", dict(order=2, scheme=MUSL), {}),
    ("musl_periodic", dict(order=2, scheme=MUSL, periodic=(1, 0, 0)), {}),
    ("musl_periodic_mixed_order", dict(order=3, n=(12, 10, 4), scheme=MUSL, periodic=(1, 0, 0)),
     {}),
    ("partialslip_walls", dict(order=3, frac=0.9, speed=2.0,
                               bcs=[(BC_PARTIALSLIP, BC_NOSLIP), (BC_SLIP, BC_PARTIALSLIP),
                                    (BC_NOSLIP, BC_SLIP)]), {"wall_vel_lo": [0, 0.2, 0] + [0] * 6}),
    ("delta_strain", dict(order=2), {"calculate_strain_based_on_delta": 1}),
    ("applied_rate", dict(order=2), {"applied_strainrate": 0.5, "applied_strainrate_time": 1.0}),
    ("external_slab", dict(order=2), {"external_loads": 1, "force_slab_lo": [0.2, 0.2, 0.2],
                                      "force_slab_hi": [0.5, 0.5, 0.4],
                                      "extforce": [1.0, -2.0, 3.0]}),
]


def apply_tweaks(s, tw):
    for k, v in tw.items():
        if isinstance(v, list):
            getattr(s, k)[:] = v
        else:
            setattr(s, k, v)


# The sorted step's tile kernels run 4 or 8 wavefronts per tile
# (MPM_OPT_TILE_WAVES; by default chosen by the particle count): both forms.
WAVES = [4, 8]


@pytest.mark.parametrize("waves", WAVES)
@pytest.mark.parametrize("name,kw,tw", CASES, ids=[c[0] for c in CASES])
def test_one_step_parity(gpu, name, kw, tw, waves):
    say(f"[parity one-step] {name}, {waves} wavefronts per tile")
    s, pos, r, idata = scenes.random_block(17, **kw)
    apply_tweaks(s, tw)
    g, o = pair(s, pos, r, idata)
    g.set_option(capi.OPT_TILE_WAVES, waves)
    g.init_internal_forces()
    o.init_internal_forces()
    compare_particles(g, o, 1e-12)
    g.step(1)
    o.step(1)
    compare_nodal(g, o, 1e-12)
    compare_particles(g, o, 1e-12)
    tg, sg, dtg = g.time()
    to, so, dto = o.time()
    assert sg == so and dtg == dto and tg == to


@pytest.mark.parametrize("waves", WAVES)
@pytest.mark.parametrize("name,kw,tw", CASES[:6] + CASES[9:11], ids=[c[0] for c in CASES[:6] + CASES[9:11]])
def test_multi_step_parity(gpu, name, kw, tw, waves):
    say(f"[parity 40 steps] {name}, {waves} wavefronts per tile")
    s, pos, r, idata = scenes.random_block(23, **kw)
    apply_tweaks(s, tw)
    g, o = pair(s, pos, r, idata)
    g.set_option(capi.OPT_TILE_WAVES, waves)
    g.init_internal_forces()
    o.init_internal_forces()
    g.step(40)
    o.step(40)
    compare_particles(g, o, 1e-10)


@pytest.mark.parametrize("name,kw,tw", [CASES[2], CASES[4], CASES[10]],
                         ids=[CASES[2][0], CASES[4][0], CASES[10][0]])
def test_unsorted_step_parity(gpu, name, kw, tw):
    """The one-thread-per-particle atomic step (MPM_OPT_SORTED = 0)."""
    say(f"[parity unsorted, 10 steps] {name}")
    s, pos, r, idata = scenes.random_block(29, **kw)
    apply_tweaks(s, tw)
    g, o = pair(s, pos, r, idata)
    g.set_option(capi.OPT_SORTED, 0)
    for ctx in (g, o):
        ctx.init_internal_forces()
        ctx.step(10)
    compare_nodal(g, o, 1e-11)
    compare_particles(g, o, 1e-11)


def test_removed_particles_are_dropped(gpu):
    """A particle leaving a non-periodic domain through a periodic-coded wall
    is removed at the next redistribution, as Redistribute does."""
    say("[parity] particle removal")
    s, pos, r, idata = scenes.random_block(31, order=2, frac=0.9, speed=0.1,
                                           bcs=[(0, 0), (BC_SLIP, BC_SLIP), (BC_SLIP, BC_SLIP)])
    r[R_VEL] += 3.0  # the block drifts through the x-hi face
    g, o = pair(s, pos, r, idata)
    for ctx in (g, o):
        ctx.init_internal_forces()
        ctx.step(25)
    assert g.num_particles() == o.num_particles() < pos.shape[1]
    compare_particles(g, o, 1e-10)


def test_individual_entry_points(gpu):
    """The per-method entry points reproduce the reference call sequence."""
    say("[parity] individual entry points")
    s, pos, r, idata = scenes.random_block(5, order=3, model=1)
    g, o = pair(s, pos, r, idata)
    o.fill_neighbors()
    for c in (g, o):
        c.reset_nodal()
        c.p2g(1, 1, 1)
    compare_nodal(g, o, 1e-12, comps=(N_MASS, N_VEL, N_VEL + 1, N_VEL + 2, N_FRC, N_FRC + 2))
    dt = o.calculate_dt()
    assert g.calculate_dt() == dt
    for c in (g, o):
        c.backup_velocity()
        c.nodal_update(dt)
        c.apply_nodal_bcs(dt, 0.0)
        c.g2p(1, 1, dt)
        c.move(dt, 1)
        c.update_volume()
        c.stress(dt, 0.0)
    compare_nodal(g, o, 1e-12)
    compare_particles(g, o, 1e-12)


def sphere_phi(s, centre_frac, radius):
    """Signed distance to a sphere on the nodes (i fastest)."""
    n = s.ncells
    dx = s.phi[0] / n[0]
    I, J, K = np.meshgrid(np.arange(n[0] + 1), np.arange(n[1] + 1), np.arange(n[2] + 1),
                          indexing="ij")
    X, Y, Z = I * dx, J * dx, K * dx
    c = np.array([centre_frac[d] * s.phi[d] for d in range(3)])
    phi = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) - radius
    return phi.transpose(2, 1, 0).ravel()


# The sorted step runs the level-set bodies and the wall UDF tables in the
# tile particle kernel (MPM_OPT_FAST = 1, the default) or in the generic
# sorted kernel (0): both against the oracle.
FAST = [1, 0]


@pytest.mark.parametrize("fast", FAST)
def test_levelset_body(gpu, fast):
    say(f"[parity] level-set sphere body, fast={fast}")
    s, pos, r, idata = scenes.random_block(9, order=2, frac=0.7, speed=1.0)
    phi = sphere_phi(s, (0.5, 0.5, 0.3), 0.12)
    g, o = pair(s, pos, r, idata)
    g.set_option(capi.OPT_FAST, fast)
    for ctx in (g, o):
        ctx.add_levelset_body(phi, BC_PARTIALSLIP, 0.3, (0.0, 0.1, 0.0))
        ctx.init_internal_forces()
        ctx.step(3)
    compare_nodal(g, o, 1e-12)
    compare_particles(g, o, 1e-11)


@pytest.mark.parametrize("fast", FAST)
def test_levelset_body_musl_order3_20_steps(gpu, fast):
    """A block falling onto a slip-wall sphere (MUSL, cubic, as the
    reference's dam break with obstacles runs): particles reach the body."""
    say(f"[parity] level-set body, MUSL order 3, 20 steps, fast={fast}")
    s, pos, r, idata = scenes.random_block(31, order=3, frac=0.6, speed=0.5, scheme=MUSL)
    r[capi.R_VEL + 1] -= 3.0
    phi = sphere_phi(s, (0.5, 0.25, 0.5), 0.1)
    g, o = pair(s, pos, r, idata)
    g.set_option(capi.OPT_FAST, fast)
    for ctx in (g, o):
        ctx.add_levelset_body(phi, BC_SLIP, 0.0)
        ctx.init_internal_forces()
    for _ in range(4):
        for ctx in (g, o):
            ctx.step(5)
        compare_particles(g, o, 1e-10)
    compare_nodal(g, o, 1e-10)


@pytest.mark.parametrize("fast", FAST)
def test_wall_udf(gpu, fast):
    say(f"[parity] wall UDF, fast={fast}")
    s, pos, r, idata = scenes.random_block(4, order=3, frac=0.95, speed=1.5)
    s.bclo[2] = BC_NOSLIP
    s.bchi[0] = BC_SLIP

    def twist(x, y, z, t):
        return (-(y - 0.3) * (1 + t), (x - 0.35) * (1 + t), 0.0)

    def shear(x, y, z, t):
        return (0.1 * z, 0.0, 0.0)

    g, o = pair(s, pos, r, idata)
    g.set_option(capi.OPT_FAST, fast)
    for ctx in (g, o):
        ctx.set_wall_udf(2, 0, twist)
        ctx.set_wall_udf(0, 1, shear)
        ctx.init_internal_forces()
        ctx.step(4)
    compare_nodal(g, o, 1e-12)
    compare_particles(g, o, 1e-11)


def test_elastic_disks_200_steps(gpu):
    say("[parity] elastic disks 200 steps")
    s, pos, r, idata = scenes.elastic_disks()
    g, o = pair(s, pos, r, idata)
    o.set_threads(16)
    for ctx in (g, o):
        ctx.init_internal_forces()
        ctx.step(200)
    pg, rg, _, _ = g.download()
    po, ro, _, _ = o.download()
    assert rel(pg, po) <= 1e-10
    assert rel(rg[R_VEL:R_VEL + 3], ro[R_VEL:R_VEL + 3]) <= 1e-10
    # The disks do not touch within 200 steps (contact at t ~ 0.59), so the
    # stress is rounding noise (~1e-14 Pa): bound its error by the modulus.
    E = 1000.0
    assert np.abs(rg[R_STRESS:R_STRESS + 6] - ro[R_STRESS:R_STRESS + 6]).max() <= 1e-10 * E
    # total mass bit-conserved
    assert rg[R_MASS].sum() == r[R_MASS].sum() and pg.shape[1] == pos.shape[1]
    eg, eo = g.energies(), o.energies()
    assert abs(eg[0] - eo[0]) <= 1e-10 * abs(eo[0])


def test_elastic_disks_musl_periodic_z_100_steps(gpu):
    """The reference's elastic-disk test runs MUSL; with z periodic the
    re-projection deposit's image contributions carry the velocity at the top
    of the step (the tile deposit's image form)."""
    say("[parity] elastic disks, MUSL, z periodic, 100 steps")
    s, pos, r, idata = scenes.elastic_disks()
    s.stress_update_scheme = MUSL
    assert s.periodic[2] == 1
    g, o = pair(s, pos, r, idata)
    o.set_threads(16)
    for ctx in (g, o):
        ctx.init_internal_forces()
        ctx.step(1)
    compare_nodal(g, o, 1e-12)
    for ctx in (g, o):
        ctx.step(99)
    pg, rg, _, _ = g.download()
    po, ro, _, _ = o.download()
    assert rel(pg, po) <= 1e-10
    assert rel(rg[R_VEL:R_VEL + 3], ro[R_VEL:R_VEL + 3]) <= 1e-10
    assert rel(rg[R_VELP:R_VELP + 3], ro[R_VELP:R_VELP + 3]) <= 1e-10
    # The disks translate rigidly until contact (t ~ 0.59), so the strain
    # rate is rounding noise: bound its error by the scale |v| / dx.
    scale = 0.2 * s.ncells[0]
    assert np.abs(rg[R_STRAINRATE:R_STRAINRATE + 6] -
                  ro[R_STRAINRATE:R_STRAINRATE + 6]).max() <= 1e-10 * scale
    assert rg[R_MASS].sum() == r[R_MASS].sum() and pg.shape[1] == pos.shape[1]


def test_elastic_disks_in_contact_200_steps(gpu):
    """Disks 0.024 apart closing at 2.83 m/s: in contact from step ~85, so
    200 steps exercise real stresses."""
    say("[parity] elastic disks in contact, 200 steps")
    s, pos, r, idata = scenes.elastic_disks(c1=(0.35, 0.35), c2=(0.65, 0.65), speed=1.0)
    g, o = pair(s, pos, r, idata)
    o.set_threads(16)
    for ctx in (g, o):
        ctx.init_internal_forces()
        ctx.step(200)
    pg, rg, _, _ = g.download()
    po, ro, _, _ = o.download()
    assert np.abs(ro[R_STRESS:R_STRESS + 6]).max() > 1.0  # real contact stresses
    assert rel(pg, po) <= 1e-10
    assert rel(rg[R_STRESS:R_STRESS + 6], ro[R_STRESS:R_STRESS + 6]) <= 1e-10
    assert rel(rg[R_VEL:R_VEL + 3], ro[R_VEL:R_VEL + 3]) <= 1e-10
    assert rg[R_MASS].sum() == r[R_MASS].sum()


def test_dam_break_small_parity(gpu):
    say("[parity] dam break 32x16x16, 60 steps")
    s, pos, r, idata = scenes.dam_break(nx=32, ny=16, nz=16)
    g, o = pair(s, pos, r, idata)
    o.set_threads(16)
    for ctx in (g, o):
        ctx.init_internal_forces()
        ctx.step(60)
    compare_particles(g, o, 1e-10)


def test_dam_break_full_size_properties(gpu):
    """Full dam break (524 288 particles): size-independent properties."""
    say("[properties] dam break full size, 30 steps")
    s, pos, r, idata = scenes.dam_break()
    g = capi.Context(s)
    g.upload(pos, r, idata)
    g.init_internal_forces()
    e0 = g.energies()
    g.step(30)
    p1, r1, _, _ = g.download()
    assert p1.shape[1] == pos.shape[1]
    assert np.isfinite(r1).all() and np.isfinite(p1).all()
    assert r1[R_MASS].sum() == r[R_MASS].sum()  # bit-conserved
    for d in range(3):
        assert p1[d].min() >= s.plo[d] and p1[d].max() <= s.phi[d]
    t, steps, _ = g.time()
    e1 = g.energies()
    # gravity impulse bounds the z-momentum change from rest
    assert e1[4] <= 0.0 and abs(e1[4]) <= e0[5] * 9.81 * t * 1.01


def rotation_partner(p0):
    """Index j of the particle at each particle's initial position rotated by
    90 degrees about the twisted column's axis x = y = 1."""
    key = lambda P: np.round(P * 2 ** 20).astype(np.int64)
    k0 = key(p0)
    idx = {tuple(k0[:, i]): i for i in range(k0.shape[1])}
    kr = key(np.stack([2.0 - p0[1], p0[0], p0[2]]))
    return np.array([idx[tuple(kr[:, i])] for i in range(kr.shape[1])])


def rotation_asymmetry(pos0, ids, rdata):
    """Relative L2 distance between the velocity field and its image under the
    90-degree rotation that maps the column's lattice onto itself."""
    perm = rotation_partner(pos0[:, ids])
    v = rdata[R_VEL:R_VEL + 3]
    vr = np.stack([-v[1], v[0], v[2]])
    return rel(v[:, perm], vr)


def test_twisted_column_small_parity(gpu):
    """Twisted column at 1/16 resolution (16 x 16 x 56 cells, 24 576
    particles, cubic): GPU against the oracle after 20 steps, and the
    90-degree rotation symmetry of the geometry kept by both (the oracle's
    own asymmetry is at rounding level)."""
    say("[parity] twisted column 16x16x56, 20 steps")
    s, pos, r, idata = scenes.twisted_column(1 / 16)
    g, o = pair(s, pos, r, idata)
    o.set_threads(16)
    for ctx in (g, o):
        ctx.init_internal_forces()
        ctx.step(20)
    compare_particles(g, o, 1e-10)
    pg, rg, _, idg = g.download()
    assert rotation_asymmetry(pos, idg, rg) <= 1e-12


def test_twisted_column_full_size_checks(gpu):
    """Twisted column at full size (256 x 256 x 896 cells, 100 663 296
    particles), 6 steps: every particle kept, the total mass unchanged, zero
    net horizontal momentum, finite energies, and the initial kinetic energy
    equal to its closed form 1/2 rho int omega(z)^2 r^2 dV = 1/2 * 1000 *
    (1/6) * 300 = 25 000 J for omega(z) = 10 sin(pi z / 12) over the
    1 x 1 x 6 column (the particle lattice is a midpoint rule of it)."""
    from exagoop_amd import column
    say("[properties] twisted column full size, 1 + 5 steps")
    out = column.run(1.0, steps=5, warmup=1, log=say)
    c = out["checks"]
    assert out["particles"] == 100663296
    assert c["particles_kept"] and c["live_particles"] == out["particles"]
    assert c["mass_rel_change"] <= 1e-14
    assert max(c["net_momentum_xy_rel"]) <= 1e-10
    assert c["energies_finite"]
    ke0, ke1 = c["kinetic_energy"]
    assert abs(ke0 - 25000.0) <= 1e-4 * 25000.0
    # the twist turns kinetic into strain energy; the total does not grow
    assert ke1 < ke0 and ke1 + c["strain_energy"] <= ke0 * (1 + 1e-6)


def test_twisted_column_quarter_size_properties(gpu):
    """Twisted column at 1/4 resolution (64 x 64 x 224 cells, 1 572 864
    particles): size-independent properties after 40 steps -- every particle
    kept, masses unchanged (the total bit-identical), the velocity field
    symmetric under the 90-degree rotation about the column axis, zero net
    horizontal momentum."""
    say("[properties] twisted column 64x64x224, 40 steps")
    s, pos, r, idata = scenes.twisted_column(1 / 4)
    g = capi.Context(s)
    g.upload(pos, r, idata)
    g.init_internal_forces()
    g.step(40)
    p1, r1, _, ids = g.download()
    assert np.array_equal(ids, np.arange(pos.shape[1]))
    assert np.isfinite(r1).all() and np.isfinite(p1).all()
    assert r1[R_MASS].sum() == r[R_MASS].sum()
    assert rotation_asymmetry(pos, ids, r1) <= 1e-10
    e = g.energies()
    scale = np.sqrt(2.0 * e[0] * e[5])
    assert abs(e[2]) <= 1e-10 * scale and abs(e[3]) <= 1e-10 * scale
