"""GPU-vs-oracle parity on the paths and lengths the core parity file leaves
out: the reference's default redistribution interval (num_redist = 10,
mpm_specs.H:89; Redistribute_Fill_Update, utilities.cpp:420-434) with
particles that cross cells, leave a free face or wrap a periodic one between
redistributions; the mixed-order domain over 40 steps; the dam break over
200 steps; wall-equivalent level-set plane and constant wall function; the
update_forces = 2 deposit; the strain energy sum.

Bars: nodal state relative L2 <= 1e-12 after one step, particle state
<= 1e-10 after 40 and 200 steps, total mass bit-identical.
"""
import sys

import numpy as np
import pytest

import oracle
from exagoop_amd import capi, scenes
from exagoop_amd.capi import (BC_NOSLIP, BC_PARTIALSLIP, BC_PERIODIC, BC_SLIP, MUSL, N_MASS,
                              N_VEL, R_MASS, R_STRESS, R_VEL, R_VELP, USL)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a, dtype=float).ravel()
    b = np.asarray(b, dtype=float).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a) if nb == 0.0 else np.linalg.norm(a - b) / nb


NODAL = (capi.N_MASS, capi.N_VEL, capi.N_VEL + 1, capi.N_VEL + 2, capi.N_FRC, capi.N_FRC + 1,
         capi.N_FRC + 2, capi.N_DVEL, capi.N_DVEL + 1, capi.N_DVEL + 2, capi.N_MASS_OLD)


def compare_nodal(g, o, tol, comps=NODAL):
    bad = {c: rel(g.nodal(c), o.nodal(c)) for c in comps if not rel(g.nodal(c), o.nodal(c)) <= tol}
    assert not bad, bad


def compare_particles(g, o, tol):
    pg, rg, ig, idg = g.download()
    po, ro, io, ido = o.download()
    assert np.array_equal(idg, ido)
    assert np.array_equal(ig, io)
    assert rel(pg, po) <= tol, rel(pg, po)
    bad = {f: rel(rg[f], ro[f]) for f in range(rg.shape[0]) if not rel(rg[f], ro[f]) <= tol}
    assert not bad, bad
    return pg, rg


def pair(s, pos, r, idata, threads=1):
    g, o = capi.Context(s), oracle.Oracle(s, threads=threads)
    g.upload(pos, r, idata)
    o.upload(pos, r, idata)
    return g, o


# ------------------------------------------------ num_redist = 10 (the default)
REDIST = [
    # name, random_block kwargs, x-velocity added to every particle
    ("walls_usl_o2", dict(order=2, frac=0.9, speed=2.5,
                          bcs=[(BC_NOSLIP, BC_SLIP), (BC_PARTIALSLIP, BC_SLIP),
                               (BC_SLIP, BC_NOSLIP)]), 0.0),
    ("walls_usl_o3", dict(order=3, frac=0.9, speed=2.5,
                          bcs=[(BC_SLIP, BC_NOSLIP), (BC_SLIP, BC_PARTIALSLIP),
                               (BC_NOSLIP, BC_SLIP)]), 0.0),
    ("walls_musl_o2", dict(order=2, frac=0.9, speed=2.5, scheme=MUSL,
                           bcs=[(BC_NOSLIP, BC_NOSLIP), (BC_SLIP, BC_SLIP),
                                (BC_PARTIALSLIP, BC_PARTIALSLIP)]), 0.0),
    ("walls_musl_o3", dict(order=3, frac=0.9, speed=2.5, scheme=MUSL,
                           bcs=[(BC_SLIP, BC_SLIP), (BC_NOSLIP, BC_NOSLIP),
                                (BC_SLIP, BC_NOSLIP)]), 0.0),
    # particles wrap the periodic x faces: unwrapped between redistributions
    ("periodic_drift_usl_o2", dict(order=2, periodic=(1, 0, 0), speed=0.5), 3.0),
    ("periodic_drift_musl_o3", dict(order=3, periodic=(1, 0, 1), speed=0.5, scheme=MUSL), 3.0),
    # particles leave through free (periodic-coded, non-periodic) x faces and
    # are removed at the next redistribution
    ("free_faces_usl_o3", dict(order=3, frac=0.9, speed=0.3,
                               bcs=[(0, 0), (BC_SLIP, BC_SLIP), (BC_SLIP, BC_SLIP)]), 3.0),
]


def redist_scene(kw, drift):
    s, pos, r, idata = scenes.random_block(41, **kw)
    s.num_redist = 10
    r[R_VEL] += drift
    return s, pos, r, idata


@pytest.mark.parametrize("name,kw,drift", REDIST, ids=[c[0] for c in REDIST])
def test_redist10_one_step(gpu, name, kw, drift):
    say(f"[parity num_redist=10, 1 step] {name}")
    s, pos, r, idata = redist_scene(kw, drift)
    g, o = pair(s, pos, r, idata)
    for c in (g, o):
        c.init_internal_forces()
        c.step(1)
    compare_nodal(g, o, 1e-12)
    compare_particles(g, o, 1e-12)


@pytest.mark.parametrize("name,kw,drift", REDIST, ids=[c[0] for c in REDIST])
def test_redist10_40_steps(gpu, name, kw, drift):
    say(f"[parity num_redist=10, 40 steps] {name}")
    s, pos, r, idata = redist_scene(kw, drift)
    g, o = pair(s, pos, r, idata)
    for c in (g, o):
        c.init_internal_forces()
    outside = 0
    for k in range(4):
        # 5 steps into the interval some particles lie outside the domain
        for c in (g, o):
            c.step(5)
        pg, _, _, _ = g.download()
        outside += int(np.sum((pg < np.array(s.plo)[:, None]) |
                              (pg > np.array(s.phi)[:, None])))
        for c in (g, o):
            c.step(5)
        compare_particles(g, o, 1e-10)
    if drift:
        assert outside > 0  # the global-deposit path of out-of-tile particles ran
    pg, rg = compare_particles(g, o, 1e-10)
    if "free" in name:
        assert pg.shape[1] < pos.shape[1]


# -------------------------------------------------------- longer horizons
def test_mixed_order_h5_40_steps(gpu):
    """z has 4 cells, so cubic falls back to linear there (mpm_init.cpp:231-232)."""
    say("[parity] mixed order (3, 3, 1), 40 steps")
    s, pos, r, idata = scenes.random_block(23, order=3, n=(12, 10, 4))
    g, o = pair(s, pos, r, idata)
    for c in (g, o):
        c.init_internal_forces()
        c.step(40)
    compare_particles(g, o, 1e-10)


def test_dam_break_64x32x32_200_steps(gpu):
    """Dam break at half resolution (65 536 particles), 200 steps of the hot
    loop (main.cpp:83-187): positions and stresses <= 1e-10, total mass
    bit-identical."""
    say("[parity] dam break 64x32x32, 200 steps, 4 and 8 wavefronts per tile")
    s, pos, r, idata = scenes.dam_break(nx=64, ny=32, nz=32)
    g, o = pair(s, pos, r, idata, threads=16)
    g8 = capi.Context(s)
    g8.upload(pos, r, idata)
    g.set_option(capi.OPT_TILE_WAVES, 4)
    g8.set_option(capi.OPT_TILE_WAVES, 8)
    for c in (g, g8, o):
        c.init_internal_forces()
    for k in range(4):
        for c in (g, g8, o):
            c.step(50)
        say(f"[parity] dam break 64x32x32: {50 * (k + 1)} steps")
    po, ro, _, _ = o.download()
    for gg in (g, g8):
        pg, rg, _, _ = gg.download()
        assert pg.shape == po.shape == pos.shape
        assert rel(pg, po) <= 1e-10, rel(pg, po)
        assert rel(rg[R_STRESS:R_STRESS + 6], ro[R_STRESS:R_STRESS + 6]) <= 1e-10
        assert rel(rg[R_VEL:R_VEL + 3], ro[R_VEL:R_VEL + 3]) <= 1e-10
        assert rg[R_MASS].sum() == r[R_MASS].sum() == ro[R_MASS].sum()
        assert gg.time()[1] == o.time()[1] == 200
        eg, eo = gg.energies(), o.energies()
        assert abs(eg[0] - eo[0]) <= 1e-10 * abs(eo[0])
        assert abs(eg[1] - eo[1]) <= 1e-10 * abs(eo[1])


# ------------------------------------------------- boundary data equivalents
# The tile particle kernel (MPM_OPT_FAST = 1, the default) and the generic
# sorted kernel (0) for level-set bodies and wall functions.
FAST = [1, 0]


@pytest.mark.parametrize("fast", FAST)
def test_levelset_plane_on_wall_nodes(gpu, fast):
    """The level-set plane whose inside is the lo-x nodes (the oracle gives
    the wall's bits with it), on the GPU against the oracle."""
    say(f"[parity] level-set plane on the wall nodes, fast={fast}")
    s, pos, r, idata = scenes.random_block(13, order=3, frac=0.8, speed=1.0,
                                           bcs=[(0, BC_SLIP), (BC_SLIP, BC_SLIP),
                                                (BC_NOSLIP, BC_SLIP)])
    n = s.ncells
    dx = s.phi[0] / n[0]
    I, _, _ = np.meshgrid(np.arange(n[0] + 1), np.arange(n[1] + 1), np.arange(n[2] + 1),
                          indexing="ij")
    phi = (I * dx - 0.5 * dx).transpose(2, 1, 0).ravel()
    g, o = pair(s, pos, r, idata)
    g.set_option(capi.OPT_FAST, fast)
    for c in (g, o):
        c.add_levelset_body(phi, BC_PARTIALSLIP, 0.3)
        c.init_internal_forces()
        c.step(1)
    compare_nodal(g, o, 1e-12)
    for c in (g, o):
        c.step(19)
    compare_particles(g, o, 1e-10)


@pytest.mark.parametrize("fast", FAST)
def test_constant_wall_function(gpu, fast):
    say(f"[parity] constant wall function, fast={fast}")
    s, pos, r, idata = scenes.random_block(14, order=2, frac=0.9, speed=1.0,
                                           bcs=[(BC_NOSLIP, BC_SLIP), (BC_SLIP, BC_SLIP),
                                                (BC_SLIP, BC_SLIP)])
    g, o = pair(s, pos, r, idata)
    g.set_option(capi.OPT_FAST, fast)
    for c in (g, o):
        c.set_wall_udf(0, 0, lambda x, y, z, t: (0.3, -0.2, 0.15))
        c.init_internal_forces()
        c.step(1)
    compare_nodal(g, o, 1e-12)
    for c in (g, o):
        c.step(19)
    compare_particles(g, o, 1e-10)


def test_p2g_update_forces_2(gpu):
    """deposit_onto_grid_momentum with update_forces = 2 deposits as 1 and
    leaves nodal component 15 zero (mpm_particle_grid_ops.cpp:398-400,
    617-625)."""
    say("[parity] p2g update_forces = 2")
    s, pos, r, idata = scenes.random_block(19, order=3, model=0)
    g, o = pair(s, pos, r, idata)
    o.fill_neighbors()
    for c in (g, o):
        c.reset_nodal()
        c.p2g(1, 1, 2)
    compare_nodal(g, o, 1e-12, comps=(N_MASS, N_VEL, N_VEL + 1, N_VEL + 2, capi.N_FRC,
                                      capi.N_FRC + 1, capi.N_FRC + 2))
    assert np.all(g.nodal(15) == 0.0) and np.all(o.nodal(15) == 0.0)


def test_strain_energy_against_oracle(gpu):
    """TKE, TSE and the momentum sums of mpm_energies against the oracle's
    Calculate_Total_Energies restatement, on disks in contact (real stresses)."""
    say("[parity] energies, elastic disks in contact, 150 steps")
    s, pos, r, idata = scenes.elastic_disks(c1=(0.35, 0.35), c2=(0.65, 0.65), speed=1.0)
    g, o = pair(s, pos, r, idata, threads=16)
    for c in (g, o):
        c.init_internal_forces()
        c.step(150)
    eg, eo = g.energies(), o.energies()
    assert eo[1] > 1e-6 * eo[0]  # in contact: strain energy present
    for q in (0, 1, 5):
        assert abs(eg[q] - eo[q]) <= 1e-10 * abs(eo[q]), (q, eg, eo)
    scale = np.sqrt(2.0 * eo[0] * eo[5])
    assert np.all(np.abs(eg[2:5] - eo[2:5]) <= 1e-10 * scale)


# ---------------------------------------------------------- caller-written nodes
def test_set_nodal_round_trip_and_uniform_field_gather(gpu):
    """mpm_set_nodal writes what mpm_get_nodal reads back, and G2P of a
    caller-written uniform nodal velocity U with zero increment gives every
    particle v' = S U with one weight sum S per particle for all three
    directions, S = 1 to the precision of the spline tables (partition of
    unity, every direction periodic), and v = alpha v + (1 - alpha) v' (the
    PIC/FLIP blend, mpm_particle_grid_ops.cpp:1030-1034)."""
    say("[nodal] set_nodal round trip, uniform field through G2P")
    s, pos, r, idata = scenes.random_block(31, order=3, model=0, periodic=(1, 1, 1), n=(8, 8, 8))
    g = capi.Context(s)
    g.upload(pos, r, idata)
    rng = np.random.default_rng(31)
    vals = rng.standard_normal(g.nn)
    g.set_nodal(capi.N_MASS_OLD, vals)
    assert np.array_equal(g.nodal(capi.N_MASS_OLD), vals)
    U = (0.3, -1.25, 2.0)
    for d in range(3):
        g.set_nodal(capi.N_VEL + d, np.full(g.nn, U[d]))
        g.set_nodal(capi.N_DVEL + d, np.zeros(g.nn))
    _, r0, _, id0 = g.download()
    g.g2p(1, 0, 1e-4)
    _, r1, _, id1 = g.download()
    assert np.array_equal(id0, id1)
    # v' = S U with S = sum of the particle's 64 weights: the cubic partition
    # of unity holds to ~7e-11 per direction (truncated constants.H literals
    # in the interpolants.H:32-170 tables), so S - 1 is ~2e-10, the same for
    # every direction.
    a = s.alpha_pic_flip
    S = r1[R_VELP] / U[0]
    assert np.max(np.abs(S - 1.0)) <= 1e-9
    for d in range(3):
        assert np.max(np.abs(r1[R_VELP + d] - S * U[d])) <= 8e-15 * abs(U[d])
        want = a * r0[R_VEL + d] + (1 - a) * r1[R_VELP + d]
        assert np.max(np.abs(r1[R_VEL + d] - want)) <= 1e-13 * (1 + np.max(np.abs(want)))
    with pytest.raises(capi.MPMError):
        g.set_nodal(99, vals)
