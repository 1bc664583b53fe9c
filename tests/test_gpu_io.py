"""GPU checks of the particle-file loader and checkpoint / restart through the
C ABI (mpm_load_particle_file, mpm_write_checkpoint, mpm_read_checkpoint).

Restart equality: a run of 10 + 10 steps and a run of 10 steps, checkpoint,
restore into a new context, Initialise_Internal_Forces (as the reference's
driver does after readCheckpointFile) and 10 steps end in the same state, to
the FP64 atomic-order spread of two uninterrupted runs (linear elastic: the
re-run initial pass leaves the stress unchanged, since it is a function of
the strain alone).
"""
import os
import sys

import numpy as np
import pytest

from exagoop_amd import capi, scenes
from exagoop_amd.capi import I_RIGID_ID, R_MASS, R_RADIUS, R_YACC

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = np.asarray(a, dtype=float).ravel()
    b = np.asarray(b, dtype=float).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a) if nb == 0.0 else np.linalg.norm(a - b) / nb


def elastic_scene():
    s, pos, r, idata = scenes.random_block(41, order=2, model=0)
    rng = np.random.default_rng(5)
    n = pos.shape[1]
    r[R_RADIUS] = rng.uniform(0.01, 0.02, n)  # cold fields: distinctive values
    r[R_YACC] = rng.uniform(-1, 1, n)
    idata[I_RIGID_ID] = rng.integers(-1, 4, n)
    return s, pos, r, idata


def run(s, pos, r, idata, n1, n2=0, ckpt=None):
    g = capi.Context(s)
    g.upload(pos, r, idata)
    g.init_internal_forces()
    g.step(n1)
    if ckpt:
        g.write_checkpoint(ckpt, output_it=7)
    if n2:
        g.step(n2)
    return g


def test_checkpoint_restart_equals_the_uninterrupted_run(gpu, tmp_path):
    say("[io] checkpoint at step 10, restart, 10 more steps")
    s, pos, r, idata = elastic_scene()
    ck = str(tmp_path / "chk00007")
    a = run(s, pos, r, idata, 10, 10, ckpt=ck)
    a2 = run(s, pos, r, idata, 20)
    lines = open(os.path.join(ck, "Header")).read().split("\n")
    assert lines[0] == "Checkpoint version: 1" and lines[1] == "10" and lines[2] == "7"
    b = capi.Context(s)
    assert b.read_checkpoint(ck) == 7
    tb0, sb0, _ = b.time()
    assert sb0 == 10
    b.init_internal_forces()
    b.step(10)
    pa, ra, ia, ida = a.download()
    pa2, ra2, _, _ = a2.download()
    pb, rb, ib, idb = b.download()
    assert np.array_equal(ida, idb) and np.array_equal(ia, ib)
    assert rel(pb, pa) <= max(1e-12, 10 * rel(pa2, pa))
    for f in range(ra.shape[0]):
        assert rel(rb[f], ra[f]) <= max(1e-12, 10 * rel(ra2[f], ra[f])), f
    # cold fields came back by upload index
    assert np.array_equal(rb[R_RADIUS], r[R_RADIUS][idb])
    assert np.array_equal(rb[R_YACC], r[R_YACC][idb])
    assert np.array_equal(ib[I_RIGID_ID], idata[I_RIGID_ID][idb])
    ta, sa, _ = a.time()
    tb, sb, _ = b.time()
    assert sa == sb == 20
    assert abs(tb - ta) <= 1e-14 * ta
    assert rb[R_MASS].sum() == r[R_MASS].sum()


def test_loaded_particle_file_steps_like_the_uploaded_arrays(gpu, tmp_path):
    say("[io] particle file -> context")
    s, pos, r, idata = scenes.random_block(43, order=2, model=1)
    path = str(tmp_path / "mpm_particles.dat")
    capi.write_particle_file(path, pos, r, idata)
    fpos, fr, fid, tot = capi.read_particle_file(path)
    a = capi.Context(s)
    ta = a.load_particle_file(path)
    assert ta == tot
    b = capi.Context(s)
    b.upload(fpos, fr, fid)
    for g in (a, b):
        g.init_internal_forces()
        g.step(5)
    pa, ra, ia, ida = a.download()
    pb, rb, ib, idb = b.download()
    assert np.array_equal(ida, idb) and np.array_equal(ia, ib)
    assert rel(pa, pb) <= 1e-12 and rel(ra[R_MASS], rb[R_MASS]) == 0.0


def trilinear(phi, s, x):
    """Nodal phi at points x (3, n) by trilinear interpolation on the nodes."""
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


def test_particles_inside_an_analytic_body_are_removed(gpu):
    """removeParticlesInsideEB with a sphere body: the removed particles are
    those where the interpolated level set is negative; none of the kept
    ones lies inside the sphere by more than the interpolation error."""
    say("[io] remove particles inside a level-set sphere")
    s, pos, r, idata = scenes.random_block(47, n=(16, 16, 16), order=2, model=0, frac=0.8)
    c, rad = (0.5, 0.5, 0.5), 0.22
    phi = capi.levelset_analytic(s, capi.LS_SPHERE, [*c, rad])
    g = capi.Context(s)
    g.upload(pos, r, idata)
    g.add_levelset_body(phi, capi.BC_NOSLIP, 0.0)
    n0 = g.num_particles()
    removed = g.remove_particles_inside_levelset()
    expect = trilinear(phi, s, pos) < 1e-20
    assert removed == int(expect.sum()) > 0
    assert g.num_particles() == n0 - removed
    p1, _, _, ids = g.download()
    assert np.array_equal(ids, np.flatnonzero(~expect))
    dist = np.sqrt(((p1 - np.array(c)[:, None]) ** 2).sum(axis=0))
    assert dist.min() >= rad - 0.5 / 16  # interpolation error near the surface
    g.init_internal_forces()
    g.step(3)
    assert g.num_particles() == n0 - removed


def test_analytic_body_bc_matches_the_oracle(gpu):
    """The level-set BC with an analytic cylinder body: GPU against the
    oracle after 10 steps."""
    import oracle
    say("[io] level-set BC with an analytic cylinder, vs oracle")
    s, pos, r, idata = scenes.random_block(53, n=(16, 12, 12), order=3, model=0, speed=1.0)
    phi = capi.levelset_analytic(s, capi.LS_CYLINDER, [0.5, 0.375, 0.375, 0.15, 0.0, 2])
    keep = trilinear(phi, s, pos) > 0.0
    pos, r, idata = pos[:, keep], r[:, keep], idata[:, keep]
    g = capi.Context(s)
    o = oracle.Oracle(s)
    for ctx in (g, o):
        ctx.upload(pos, r, idata)
        ctx.add_levelset_body(phi, capi.BC_SLIP, 0.0)
        ctx.init_internal_forces()
        ctx.step(10)
    pg, rg, ig, idg = g.download()
    po, ro, io_, ido = o.download()
    assert np.array_equal(idg, ido)
    assert rel(pg, po) <= 1e-10
    for f in range(rg.shape[0]):
        assert rel(rg[f], ro[f]) <= 1e-10, f


def test_ascii_particle_output(gpu, tmp_path):
    say("[io] ASCII particle output")
    s, pos, r, idata = scenes.random_block(59, order=2, model=0)
    g = capi.Context(s)
    g.upload(pos, r, idata)
    g.init_internal_forces()
    g.step(3)
    t = g.time()[0]
    path = g.write_ascii_particles(str(tmp_path / "matpnt"), t)
    assert path == str(tmp_path / "matpnt") + "_t%.6f" % t
    lines = open(path).read().split("\n")
    n = int(lines[0])
    p, rr, ii, ids = g.download()
    assert n == p.shape[1]
    rows = np.array([[float(v) for v in l.split()] for l in lines[1:1 + n]])
    assert rows.shape == (n, 3 + 2 + capi.NREAL + capi.NINT)
    assert np.array_equal(rows[:, :3].T, p)  # 17 digits round-trip exactly
    assert np.array_equal(rows[:, 3].astype(np.int64), ids)
    assert np.array_equal(rows[:, 5:5 + capi.NREAL].T, rr)
    assert np.array_equal(rows[:, 5 + capi.NREAL:].T.astype(np.int32), ii)


def test_checkpoint_carries_the_thermal_fields(gpu, tmp_path):
    """With the thermal module on, the checkpoint carries the thermal record
    by upload index; a context with the module enabled reads it back, and
    one without it reads the mechanical record only."""
    say("[io] thermal checkpoint")
    s, pos, r, idata = elastic_scene()
    n = pos.shape[1]
    rng = np.random.default_rng(11)
    th = np.zeros((capi.NTHERMAL, n))
    th[capi.TH_TEMPERATURE] = rng.uniform(0, 1, n)
    th[capi.TH_SPECIFIC_HEAT] = 1.0
    th[capi.TH_CONDUCTIVITY] = 30.0
    a = capi.Context(s)
    a.upload(pos, r, idata)
    a.enable_thermal()
    a.upload_thermal(th)
    a.init_internal_forces()
    a.step(5)
    ck = str(tmp_path / "chk_thermal")
    a.write_checkpoint(ck)
    b = capi.Context(s)
    b.enable_thermal()
    b.read_checkpoint(ck)
    tha, thb = a.download_thermal(), b.download_thermal(n)
    assert np.array_equal(np.isnan(tha), np.isnan(thb))
    live = ~np.isnan(tha)
    assert np.array_equal(tha[live], thb[live])
    c = capi.Context(s)
    c.read_checkpoint(ck)
    pc, _, _, _ = c.download()
    pa, _, _, _ = a.download()
    assert np.array_equal(pc, pa)


def test_checkpoint_header_of_the_eb_build(gpu, tmp_path):
    """The reference's default build (EB on) writes EB_generate_max_level
    before the time (mpm_particle_outputs.cpp:381-385); the reader takes the
    time from the last line in either layout."""
    say("[io] checkpoint Header with the EB line")
    s, pos, r, idata = elastic_scene()
    ck = str(tmp_path / "chk_eb")
    a = run(s, pos, r, idata, 7, ckpt=ck)
    ta, sa, _ = a.time()
    hp = os.path.join(ck, "Header")
    lines = open(hp).read().split("\n")
    assert float(lines[3]) == ta and ta > 0.0
    open(hp, "w").write("\n".join(lines[:3] + ["0"] + lines[3:]))
    b = capi.Context(s)
    assert b.read_checkpoint(ck) == 7
    tb, sb, _ = b.time()
    assert sb == sa == 7 and tb == ta


def test_thermal_context_needs_thermal_fields_in_the_checkpoint(gpu, tmp_path):
    say("[io] thermal context, checkpoint without thermal.bin")
    s, pos, r, idata = elastic_scene()
    ck = str(tmp_path / "chk_mech")
    run(s, pos, r, idata, 3, ckpt=ck)
    b = capi.Context(s)
    b.enable_thermal()
    with pytest.raises(capi.MPMError, match="thermal"):
        b.read_checkpoint(ck)


def test_thermal_particle_file_loads_both_records(gpu, tmp_path):
    say("[io] thermal particle file -> context")
    s, pos, r, idata = scenes.random_block(43, order=2, model=1)
    n = pos.shape[1]
    rng = np.random.default_rng(12)
    th = np.zeros((capi.NTHERMAL, n))
    th[capi.TH_TEMPERATURE] = rng.uniform(0, 1, n)
    th[capi.TH_SPECIFIC_HEAT] = rng.uniform(1, 2, n)
    th[capi.TH_CONDUCTIVITY] = rng.uniform(20, 40, n)
    th[capi.TH_HEAT_SOURCE] = rng.uniform(-10, 10, n)
    path = str(tmp_path / "mpm_particles_thermal.dat")
    capi.write_particle_file(path, pos, r, idata, th)
    fpos, fr, fid, fth, tot = capi.read_particle_file(path, thermal=True)
    a = capi.Context(s)
    a.enable_thermal()
    assert a.load_particle_file(path, thermal=True) == tot
    b = capi.Context(s)
    b.upload(fpos, fr, fid)
    b.enable_thermal()
    b.upload_thermal(fth)
    for g in (a, b):
        g.init_internal_forces()
        g.step(5)
    ta, tb = a.download_thermal(), b.download_thermal()
    for f in range(capi.NTHERMAL):
        assert rel(ta[f], tb[f]) <= 1e-12, f
