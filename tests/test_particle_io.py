"""Particle input formats (CPU): the library's ASCII reader, in-built
generator and writer (mpm_read_particle_file, mpm_autogen_particles,
mpm_write_particle_file) against the oracle's restatement
(oracle/particle_io.py), and the oracle pinned on the reference's own
particle file (Tests/Dam_Break_With_Obstacles/mpm_particles_80.dat, its first
records in tests/golden/particle_files) and on closed forms."""
import math
import os

import numpy as np
import pytest

from exagoop_amd import capi
from oracle import particle_io as oio

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "particle_files")
HEAD = os.path.join(GOLD, "dam_break_obstacles_head.dat")
EMPTY = os.path.join(GOLD, "dam_break_obstacles_stl.dat")


def test_oracle_reads_the_reference_file_as_the_2d_build():
    pos, r, idata, (tm, tv) = oio.read_particle_file(HEAD, dim=2, max_records=5)
    rows = [l.split("\t") for l in open(HEAD).read().split("\n")[3:8]]
    assert pos.shape == (2, 5)
    for i, row in enumerate(rows):
        phase, x, y, rad, dens, vx, vy, cm, K, gam, mu = row
        assert idata[oio.I_PHASE, i] == int(phase) == 0
        assert (pos[0, i], pos[1, i]) == (float(x), float(y))
        assert r[oio.R_RADIUS, i] == float(rad) and r[oio.R_DENSITY, i] == float(dens)
        assert (r[oio.R_VEL, i], r[oio.R_VEL + 1, i]) == (float(vx), float(vy))
        assert idata[oio.I_MODEL, i] == int(cm) == 1
        assert (r[oio.R_BULK, i], r[oio.R_GAMMA, i], r[oio.R_VISC, i]) == \
            (float(K), float(gam), float(mu))
        assert r[oio.R_E, i] == 0.0 and r[oio.R_NU, i] == 0.0
        # volume with constants.H's literals, not math.pi
        vol = 1.3333333333 * 3.141592654 * float(rad) ** 3
        assert r[oio.R_VOLUME, i] == vol and r[oio.R_VOLINIT, i] == vol
        assert r[oio.R_MASS, i] == float(dens) * vol
        assert vol != 4.0 / 3.0 * math.pi * float(rad) ** 3
    assert abs(tm - r[oio.R_MASS].sum()) <= 1e-15 * tm


def test_the_reference_files_are_rejected_by_the_3d_build():
    for path in (HEAD, EMPTY):
        with pytest.raises(oio.FileError, match="Dimension mismatch"):
            oio.read_particle_file(path)
        with pytest.raises(capi.MPMError, match="Dimension mismatch"):
            capi.read_particle_file(path)


def write_text(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


@pytest.mark.parametrize("text,msg", [
    ("dim: 3\nnumber_of_material_points: 0\n# x\n", "Invalid number_of_material_points"),
    ("dims: 3\nnumber_of_material_points: 2\n# x\n", "Expected 'dim:'"),
    ("dim: 3\nnumber_of_points: 2\n# x\n", "Expected 'number_of_material_points:'"),
    ("dim: 3\nnumber_of_material_points: 1\nx y\n", "Expected header line"),
    # a trailing newline is not EOF: the next extraction fails instead
    ("dim: 3\nnumber_of_material_points: 2\n# h\n0 0.1 0.1 0.1 0.01 1000 0 0 0 0 1e4 0.3\n",
     "Error reading phase"),
    ("dim: 3\nnumber_of_material_points: 2\n# h\n0 0.1 0.1 0.1 0.01 1000 0 0 0 0 1e4 0.3",
     "Unexpected EOF"),
    ("dim: 3\nnumber_of_material_points: 1\n# h\n0 0.1 0.1 0.1 0.01 1000 0 0 0 1 2e4 7\n",
     "Error reading Dynamic_viscosity"),
    ("dim: 3\nnumber_of_material_points: 1\n# h\n0 0.1 0.1 0.1 0.01 1000 0 0 0 3 1 2\n",
     "Incorrect constitutive model"),
])
def test_abort_conditions_are_errors(tmp_path, text, msg):
    path = write_text(tmp_path, "p.dat", text)
    with pytest.raises(oio.FileError, match=msg):
        oio.read_particle_file(path)
    with pytest.raises(capi.MPMError, match=msg):
        capi.read_particle_file(path)


def random_particles(seed, n=300):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, 1, (3, n))
    r = np.zeros((oio.NREAL, n))
    idata = np.zeros((oio.NINT, n), dtype=np.int32)
    idata[oio.I_MODEL] = rng.integers(0, 3, n)
    idata[oio.I_MODEL, 0] = 1
    idata[oio.I_PHASE] = (rng.uniform(size=n) < 0.1).astype(np.int32)
    r[oio.R_RADIUS] = rng.uniform(1e-3, 1e-2, n)
    r[oio.R_DENSITY] = rng.uniform(500, 3000, n)
    r[oio.R_VEL:oio.R_VEL + 3] = rng.normal(size=(3, n))
    r[oio.R_E], r[oio.R_NU] = rng.uniform(1e3, 1e7, n), rng.uniform(0.1, 0.45, n)
    r[oio.R_BULK], r[oio.R_GAMMA], r[oio.R_VISC] = rng.uniform(1e4, 1e6, n), 7.0, 1e-3
    return pos, r, idata


def test_writer_matches_the_generator_layout_and_round_trips(tmp_path):
    pos, r, idata = random_particles(3)
    a, b = str(tmp_path / "lib.dat"), str(tmp_path / "orc.dat")
    capi.write_particle_file(a, pos, r, idata)
    oio.write_particle_file(b, pos, r, idata)
    assert open(a).read() == open(b).read()  # byte for byte
    lines = open(a).read().split("\n")
    assert lines[0] == "dim: 3" and lines[1] == "number_of_material_points: 300"
    assert lines[2] == ("# phase x y z radius density vx vy vz cm_id Bulk_modulus "
                        "Gamma_pressure Dynamic_viscosity")
    pl, rl, il, tl = capi.read_particle_file(a)
    po, ro, io_, to = oio.read_particle_file(a)
    assert np.array_equal(pl, po) and np.array_equal(rl, ro) and np.array_equal(il, io_)
    assert tl == to
    # values are the inputs rounded to 7 significant digits
    assert np.array_equal(pl, np.vectorize(lambda v: float(f"{v:.6e}"))(pos))
    m = idata[oio.I_MODEL]
    assert np.array_equal(rl[oio.R_E][m != 1], np.vectorize(lambda v: float(f"{v:.6e}"))(
        r[oio.R_E][m != 1]))
    assert np.all(rl[oio.R_E][m == 1] == 0) and np.all(rl[oio.R_BULK][m != 1] == 0)
    # totals over phase-0 particles
    ph0 = il[oio.I_PHASE] == 0
    assert abs(tl[0] - rl[oio.R_MASS][ph0].sum()) <= 1e-12 * tl[0]


@pytest.mark.parametrize("multi,ppc", [(0, (1, 1, 1)), (1, (2, 2, 2)), (1, (3, 2, 4))])
def test_autogen_matches_the_oracle(multi, ppc):
    s = capi.default_specs()
    s.plo[:] = [0.0, -0.5, 0.25]
    s.phi[:] = [1.0, 0.5, 1.25]
    s.ncells[:] = [10, 8, 6]
    args = dict(vel=[0.1, -0.2, 0.3], ppc=ppc, dens=1200.0, model=0, E=2e5, nu=0.25,
                multi_per_cell=multi)
    lo, hi = [0.2, -0.3, 0.3], [0.71, 0.4, 1.0]
    pl, rl, il, tl = capi.autogen_particles(s, lo, hi, **args)
    po, ro, io_, to = oio.autogen_particles(list(s.plo), list(s.phi), list(s.ncells), lo, hi,
                                            **args)
    assert pl.shape == po.shape and pl.shape[1] > 0
    assert np.array_equal(pl, po) and np.array_equal(rl, ro) and np.array_equal(il, io_)
    assert tl == pytest.approx(to, rel=1e-15)
    if multi:
        # the reference's 3-D build puts ppc_x ppc_y particles per cell, all at
        # the cell's mid-height, each with 1/(ppc_x ppc_y) of the cell volume
        dz = (s.phi[2] - s.plo[2]) / s.ncells[2]
        frac = (pl[2] - s.plo[2]) / dz
        assert np.allclose(frac - np.floor(frac), 0.5)
        cell = 0.1 * 0.125 * (1.0 / 6)
        assert np.allclose(rl[oio.R_VOLUME], cell / (ppc[0] * ppc[1]))
    rad = (3.0 * 0.25 * rl[oio.R_VOLUME][0] / 3.141592654) ** (1.0 / 3.0)
    assert rl[oio.R_RADIUS][0] == rad


def test_thermal_columns_write_and_read_as_the_use_temp_build(tmp_path):
    """The generator's temperature columns (T spheat thermcond heatsrc after
    the model's) written byte for byte as the restated generator writes them,
    and read back as the USE_TEMP reader does (heat flux zero)."""
    pos, r, idata = random_particles(5)
    n = pos.shape[1]
    rng = np.random.default_rng(6)
    th = np.zeros((capi.NTHERMAL, n))
    th[capi.TH_TEMPERATURE] = rng.uniform(250, 400, n)
    th[capi.TH_SPECIFIC_HEAT] = rng.uniform(100, 4000, n)
    th[capi.TH_CONDUCTIVITY] = rng.uniform(0.1, 400, n)
    th[capi.TH_HEAT_FLUX:capi.TH_HEAT_FLUX + 3] = 9.0  # not written
    th[capi.TH_HEAT_SOURCE] = rng.uniform(-1e3, 1e3, n)
    a, b = str(tmp_path / "lib.dat"), str(tmp_path / "orc.dat")
    capi.write_particle_file(a, pos, r, idata, th)
    oio.write_particle_file(b, pos, r, idata, th)
    assert open(a).read() == open(b).read()
    assert open(a).read().split("\n")[2].endswith("Dynamic_viscosity T spheat thermcond heatsrc")
    pl, rl, il, thl, tl = capi.read_particle_file(a, thermal=True)
    po, ro, io_, tho, to = oio.read_particle_file(a, thermal=True)
    assert np.array_equal(pl, po) and np.array_equal(rl, ro) and np.array_equal(il, io_)
    assert np.array_equal(thl, tho) and tl == to
    six = np.vectorize(lambda v: float(f"{v:.6e}"))
    for f in (capi.TH_TEMPERATURE, capi.TH_SPECIFIC_HEAT, capi.TH_CONDUCTIVITY,
              capi.TH_HEAT_SOURCE):
        assert np.array_equal(thl[f], six(th[f]))
    assert np.all(thl[capi.TH_HEAT_FLUX:capi.TH_HEAT_FLUX + 3] == 0.0)
    # the mechanical reader stops short of the extra columns: the next
    # record's phase is read from the temperature column and the record
    # after fails
    with pytest.raises(capi.MPMError):
        capi.read_particle_file(a)


@pytest.mark.parametrize("cut,msg", [(1, "heat_source"), (2, "thermal_conductivity"),
                                     (3, "specific_heat"), (4, "temperature")])
def test_truncated_thermal_columns_are_errors(tmp_path, cut, msg):
    row = "0 0.1 0.1 0.1 0.01 1000 0 0 0 0 1e4 0.3 300 1000 2.5 7".split()
    text = "dim: 3\nnumber_of_material_points: 1\n# h\n" + " ".join(row[:-cut]) + "\n"
    path = write_text(tmp_path, "t.dat", text)
    with pytest.raises(oio.FileError, match="Error reading " + msg):
        oio.read_particle_file(path, thermal=True)
    with pytest.raises(capi.MPMError, match="Error reading " + msg):
        capi.read_particle_file(path, thermal=True)
