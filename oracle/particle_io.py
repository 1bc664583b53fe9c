"""Particle input formats, restated in Python -- TEST INFRASTRUCTURE ONLY.

The checker of libexagoop_amd's mpm_read_particle_file,
mpm_autogen_particles and mpm_write_particle_file; only tests/ import it.

* read_particle_file: InitParticles(filename) (Source/mpm_init.cpp:942-1167),
  written for any dimension as the reference's AMREX_SPACEDIM code is (dim is
  the build's dimension; the library is the 3-D build).
* autogen_particles: InitParticles(mincoords, maxcoords, ...) and
  generate_particle (mpm_init.cpp:1199-1382, 1386-1473), 3-D build.
* write_particle_file: write_particles_ascii_streaming
  (Tests/*/PreProcess/Generate_MPs_Inputfile_Generic.py:1020-1127).

Pins (tests/test_particle_io.py): the reference's own particle file
(Tests/Dam_Break_With_Obstacles/mpm_particles_80.dat, its first records
copied to tests/golden/particle_files) parsed as the 2-D build does, with
every value compared with the file's text; the 0-particle file; closed forms
of volume, mass and radius with the constants.H literals.
"""
import math
import re

import numpy as np

NREAL, NINT = 46, 3
R_RADIUS, R_VEL, R_DEFGRAD, R_VOLUME, R_MASS, R_DENSITY = 0, 1, 25, 34, 35, 36
R_JACOBIAN, R_PRESSURE, R_VOLINIT, R_E, R_NU, R_BULK, R_GAMMA, R_VISC = 37, 38, 39, 40, 41, 42, 43, 44
I_PHASE, I_RIGID_ID, I_MODEL = 0, 1, 2

FOURBYTHREE = 1.3333333333  # constants.H:10
PI = 3.141592654            # constants.H:12
THREE = 3.0                 # constants.H:13
FOURTH = 0.25               # constants.H:21


class FileError(ValueError):
    """A condition InitParticles aborts on."""


def _blank(n, dim):
    pos = np.zeros((dim, n))
    r = np.zeros((NREAL, n))
    idata = np.zeros((NINT, n), dtype=np.int32)
    for d in range(dim):
        r[R_DEFGRAD + d * dim + d] = 1.0  # deformation_gradient diagonal (mpm_init.cpp:1124-1131)
    r[R_JACOBIAN] = 1.0
    idata[I_RIGID_ID] = -1
    return pos, r, idata


def read_particle_file(path, dim=3, max_records=None, thermal=False):
    """(pos, rdata, idata, (total mass, total volume)) as InitParticles fills
    the particle tile; raises FileError where it aborts.  max_records reads
    only the first records (for a file cut after its head).  thermal: the
    USE_TEMP build's layout (mpm_init.cpp:1097-1108), returning
    (pos, rdata, idata, th, totals) with th rows T, c_p, k, q (3, zero), Q."""
    with open(path) as f:
        text = f.read()
    lines = text.split("\n")
    head = (lines[0] + " " + (lines[1] if len(lines) > 1 else "")).split()
    if len(head) < 4:
        raise FileError("Error reading header")
    if head[0] != "dim:":
        raise FileError("mpm_particles.dat: Expected 'dim:' at line 1")
    if int(head[1]) != dim:
        raise FileError("Dimension mismatch between particle file and build")
    if head[2] != "number_of_material_points:":
        raise FileError("Expected 'number_of_material_points:' at line 2")
    n = int(head[3])
    if n <= 0:
        raise FileError("Invalid number_of_material_points")
    if len(lines) < 3 or not lines[2].startswith("#"):
        raise FileError("Expected header line beginning with '#'")
    body = "\n".join(lines[3:])
    toks = [(m.group(0), m.end()) for m in re.finditer(r"\S+", body)]
    if max_records is not None:
        n = min(n, max_records)
    pos, r, idata = _blank(n, dim)
    th = np.zeros((7, n))
    k = 0
    at = 0  # characters consumed (the stream position)
    tmass = tvol = 0.0

    def take(conv, what):
        # safe_read: a failed extraction aborts with the field's message
        nonlocal k, at
        if k >= len(toks):
            raise FileError("Error reading " + what)
        v = conv(toks[k][0])
        at = toks[k][1]
        k += 1
        return v

    for i in range(n):
        # ifs.peek() == EOF only when nothing at all follows the last value
        if at >= len(body):
            raise FileError("Unexpected EOF while reading particle data")
        phase = take(int, "phase")
        idata[I_PHASE, i] = phase
        for d in range(dim):
            pos[d, i] = take(float, "particle position")
        rad = take(float, "radius")
        dens = take(float, "density")
        r[R_RADIUS, i] = rad
        r[R_DENSITY, i] = dens
        for d in range(dim):
            r[R_VEL + d, i] = take(float, "velocity")
        model = take(int, "constitutive_model")
        idata[I_MODEL, i] = model
        if model in (0, 2):
            r[R_E, i] = take(float, "E")
            r[R_NU, i] = take(float, "nu")
        elif model == 1:
            r[R_BULK, i] = take(float, "Bulk_modulus")
            r[R_GAMMA, i] = take(float, "Gama_pressure")
            r[R_VISC, i] = take(float, "Dynamic_viscosity")
        else:
            raise FileError("Incorrect constitutive model")
        if thermal:
            th[0, i] = take(float, "temperature")
            th[1, i] = take(float, "specific_heat")
            th[2, i] = take(float, "thermal_conductivity")
            th[6, i] = take(float, "heat_source")
        vol = FOURBYTHREE * PI * math.pow(rad, THREE)
        r[R_VOLUME, i] = vol
        r[R_MASS, i] = dens * vol
        r[R_VOLINIT, i] = vol
        if phase == 0:
            tmass += dens * vol
            tvol += vol
    if thermal:
        return pos, r, idata, th, (tmass, tvol)
    return pos, r, idata, (tmass, tvol)


def autogen_particles(plo, phi, ncells, mincoords, maxcoords, vel, ppc, dens, model, E=0.0,
                      nu=0.0, K=0.0, gamma=0.0, visc=0.0, multi_per_cell=1):
    """The in-built generator of the 3-D build (one box, cells x fastest)."""
    dx = [(phi[d] - plo[d]) / ncells[d] for d in range(3)]
    cell_vol = dx[0] * dx[1] * dx[2]
    # ppc_z stays 1 in 3-D: the if / else-if chain of mpm_init.cpp:1239-1251
    # takes the SPACEDIM >= 2 branch and never reaches SPACEDIM == 3.
    ppc_x, ppc_y, ppc_z = ppc[0], ppc[1], 1
    offx = [(i + 0.5) / ppc_x for i in range(ppc_x)]
    offy = [(i + 0.5) / ppc_y for i in range(ppc_y)]
    offz = [(i + 0.5) / ppc_z for i in range(ppc_z)]
    vol = cell_vol / (ppc_x * ppc_y * ppc_z) if multi_per_cell else cell_vol
    pts = []
    for k in range(ncells[2]):
        for j in range(ncells[1]):
            for i in range(ncells[0]):
                iv = (i, j, k)
                if not multi_per_cell:
                    c = [plo[d] + (iv[d] + 0.5) * dx[d] for d in range(3)]
                    cand = [c]
                else:
                    base = [plo[d] + iv[d] * dx[d] for d in range(3)]
                    cand = [[base[0] + ox * dx[0], base[1] + oy * dx[1], base[2] + oz * dx[2]]
                            for ox in offx for oy in offy for oz in offz]
                for c in cand:
                    if all(mincoords[d] <= c[d] <= maxcoords[d] for d in range(3)):
                        pts.append(c)
    n = len(pts)
    pos, r, idata = _blank(n, 3)
    if n:
        pos[:] = np.asarray(pts).T
    rad = math.pow(THREE * FOURTH * vol / PI, 1.0 / 3.0)
    r[R_RADIUS] = rad
    r[R_DENSITY] = dens
    for d in range(3):
        r[R_VEL + d] = vel[d]
    idata[I_MODEL] = model
    r[R_E], r[R_NU], r[R_BULK], r[R_GAMMA], r[R_VISC] = E, nu, K, gamma, visc
    r[R_VOLUME] = vol
    r[R_MASS] = dens * vol
    r[R_VOLINIT] = vol
    tm = tv = 0.0
    for _ in range(n):
        tm += dens * vol
        tv += vol
    return pos, r, idata, (tm, tv)


def write_particle_file(path, pos, r, idata, th=None):
    """The generator's ASCII layout for a 3-D file: header, '#' column line
    (core fields, then the model fields of the first particle's model, then
    with th the temperature fields T spheat thermcond heatsrc), rows
    tab-separated with integers as str() and reals as f"{v:.6e}"."""
    n = pos.shape[1]
    m0 = int(idata[I_MODEL, 0])
    cols = ["phase", "x", "y", "z", "radius", "density", "vx", "vy", "vz", "cm_id"]
    cols += ["Bulk_modulus", "Gamma_pressure", "Dynamic_viscosity"] if m0 == 1 else ["E", "nu"]
    if th is not None:
        cols += ["T", "spheat", "thermcond", "heatsrc"]
    out = ["dim: 3", f"number_of_material_points: {n}", "# " + " ".join(cols)]
    for i in range(n):
        model = int(idata[I_MODEL, i])
        row = [str(int(idata[I_PHASE, i]))]
        row += [f"{pos[d, i]:.6e}" for d in range(3)]
        row += [f"{r[R_RADIUS, i]:.6e}", f"{r[R_DENSITY, i]:.6e}"]
        row += [f"{r[R_VEL + d, i]:.6e}" for d in range(3)]
        row.append(str(model))
        if model == 1:
            row += [f"{r[f, i]:.6e}" for f in (R_BULK, R_GAMMA, R_VISC)]
        else:
            row += [f"{r[f, i]:.6e}" for f in (R_E, R_NU)]
        if th is not None:
            row += [f"{th[f, i]:.6e}" for f in (0, 1, 2, 6)]
        out.append("\t".join(row))
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
