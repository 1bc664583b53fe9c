"""Scene builders for the benchmark and test workloads.

Particles are laid out as the reference's generator does
(Tests/*/PreProcess/Generate_MPs_Inputfile_Generic.py:179-181, 241): N
particles per cell per direction at offsets (2i-1)/(2N), inclusive shape
tests, radius (3V/4pi)^(1/3) of the per-particle volume.  The particle
record is initialised as InitParticles does (mpm_init.cpp:1087-1140):
V = (4/3) pi r^3 with the reference's truncated constants, m = rho V,
J = 1, V0 = V, F = I, strain/stress/strain-rate zero.
"""
import math

import numpy as np

from .capi import (NREAL, NINT, R_RADIUS, R_VEL, R_DEFGRAD, R_VOLUME, R_MASS, R_DENSITY,
                   R_JACOBIAN, R_VOLINIT, R_E, R_NU, R_BULK, R_GAMMA, R_VISC, I_PHASE,
                   I_RIGID_ID, I_MODEL, BC_NOSLIP, BC_PERIODIC, BC_SLIP, USL, default_specs)

FOURBYTHREE = 1.3333333333  # constants.H:10
PI_REF = 3.141592654        # constants.H:12


def lattice(specs, ppc, inside, box=None):
    """Cell-offset lattice of all cells (or of the cells overlapping
    box = ((lo, hi),) * 3), filtered by inside(x, y, z)."""
    n = specs.ncells
    plo, phi = specs.plo, specs.phi
    dx = [(phi[d] - plo[d]) / n[d] for d in range(3)]
    offs = [(2 * np.arange(1, ppc[d] + 1) - 1) / (2 * ppc[d]) for d in range(3)]
    coords = []
    for d in range(3):
        c0, c1 = 0, n[d]
        if box is not None:
            c0 = max(0, int(np.floor((box[d][0] - plo[d]) / dx[d])) - 1)
            c1 = min(n[d], int(np.ceil((box[d][1] - plo[d]) / dx[d])) + 1)
        c = (np.arange(c0, c1)[:, None] + offs[d][None, :]).ravel() * dx[d] + plo[d]
        coords.append(c)
    X, Y, Z = np.meshgrid(coords[0], coords[1], coords[2], indexing="ij")
    X, Y, Z = X.ravel(), Y.ravel(), Z.ravel()
    keep = inside(X, Y, Z)
    pos = np.stack([X[keep], Y[keep], Z[keep]])
    vol_particle = dx[0] * dx[1] * dx[2] / (ppc[0] * ppc[1] * ppc[2])
    return pos, vol_particle


def make_records(pos, vol_particle, density, model, E=0.0, nu=0.0, K=0.0, gamma=0.0, mu=0.0,
                 vel=None):
    np_ = pos.shape[1]
    r = np.zeros((NREAL, np_))
    idata = np.zeros((NINT, np_), dtype=np.int32)
    rad = (3.0 / 4.0 * vol_particle / np.pi) ** (1.0 / 3.0)
    r[R_RADIUS] = rad
    vol = FOURBYTHREE * PI_REF * rad ** 3.0
    r[R_VOLUME] = vol
    r[R_DENSITY] = density
    r[R_MASS] = density * vol
    r[R_JACOBIAN] = 1.0
    r[R_VOLINIT] = vol
    for d in range(3):
        r[R_DEFGRAD + 4 * d] = 1.0
    if model == 1:
        r[R_BULK], r[R_GAMMA], r[R_VISC] = K, gamma, mu
    else:
        r[R_E], r[R_NU] = E, nu
    if vel is not None:
        r[R_VEL:R_VEL + 3] = vel
    idata[I_PHASE] = 0
    idata[I_RIGID_ID] = -1
    idata[I_MODEL] = model
    return r, idata


def elastic_disks(nx=64, ny=64, nz=4, ppc=2, c1=(0.3, 0.3), c2=(0.7, 0.7), speed=0.1):
    """Two linear-elastic disks colliding, extruded 3-D: unit square,
    64x64x4 cells (dx = 1/64), z periodic, 2x2x2 particles per cell, r = 0.2
    at (0.3,0.3) and (0.7,0.7), v = +-(0.1,0.1,0), E = 1000, nu = 0.3,
    rho = 1000, quadratic B-splines, fixed dt = 1e-4.  With these values the
    disks first touch at t ~ 0.59 (gap 0.166 closing at 0.283)."""
    s = default_specs()
    s.plo[:] = [0.0, 0.0, 0.0]
    s.phi[:] = [1.0, ny / nx, nz / nx]
    s.ncells[:] = [nx, ny, nz]
    s.periodic[:] = [0, 0, 1]
    s.bclo[2] = s.bchi[2] = BC_PERIODIC
    s.order_scheme = 2
    s.num_redist = 1
    s.gravity[:] = [0.0, 0.0, 0.0]
    s.fixed_timestep = 1
    s.timestep = 1e-4
    s.stress_update_scheme = USL

    def inside(X, Y, Z):
        a = (X - c1[0]) ** 2 + (Y - c1[1]) ** 2 <= 0.2 ** 2
        b = (X - c2[0]) ** 2 + (Y - c2[1]) ** 2 <= 0.2 ** 2
        return a | b

    pos, vp = lattice(s, (ppc, ppc, ppc), inside)
    vel = np.zeros((3, pos.shape[1]))
    first = (pos[0] - c1[0]) ** 2 + (pos[1] - c1[1]) ** 2 <= 0.2 ** 2
    vel[0] = np.where(first, speed, -speed)
    vel[1] = np.where(first, speed, -speed)
    r, idata = make_records(pos, vp, 1000.0, 0, E=1000.0, nu=0.3, vel=vel)
    return s, pos, r, idata


def dam_break(nx=128, ny=64, nz=64, width=0.5, height=0.5, ppc=2, order=3, scheme=USL,
              alpha=0.9):
    """Compressible-fluid dam break: domain 2x1x1, 128x64x64 cells
    (dx = 1/64), water block [0,width]x[0,1]x[0,height], 2x2x2 particles per
    cell, Tait EOS K = 2e6, gamma = 7, mu = 1e-3, rho = 1000, gravity -9.81 z,
    no-slip walls, cubic B-splines, CFL 0.3."""
    s = default_specs()
    s.plo[:] = [0.0, 0.0, 0.0]
    s.phi[:] = [2.0, 2.0 * ny / nx, 2.0 * nz / nx]
    s.ncells[:] = [nx, ny, nz]
    s.periodic[:] = [0, 0, 0]
    for d in range(3):
        s.bclo[d] = s.bchi[d] = BC_NOSLIP
    s.order_scheme = order
    s.num_redist = 1
    s.gravity[:] = [0.0, 0.0, -9.81]
    s.fixed_timestep = 0
    s.CFL = 0.3
    s.alpha_pic_flip = alpha
    s.stress_update_scheme = scheme

    def inside(X, Y, Z):
        return (X <= width) & (Z <= height)

    pos, vp = lattice(s, (ppc, ppc, ppc), inside)
    r, idata = make_records(pos, vp, 1000.0, 1, K=2e6, gamma=7.0, mu=1e-3)
    return s, pos, r, idata


def twisted_column(res=1.0, ppc=2, omega0=10.0):
    """Linear-elastic column [0.5, 1.5]^2 x [0, 6] in the domain [0, 2]^2 x
    [0, 7], clamped base (no-slip z-lo, slip elsewhere), initial twist
    v = w(z) z x (x-1, y-1, 0) with w(z) = omega0 sin(pi z / 12), E = 1e6,
    nu = 0.3, rho = 1000, cubic B-splines, CFL 0.3, USL.  res = 1 is the full
    size: 256 x 256 x 896 cells (dx = 1/128), 2 x 2 x 2 particles per cell,
    100 663 296 particles; res scales the cell count per direction (the
    geometry stays)."""
    n = [int(round(256 * res)), int(round(256 * res)), int(round(896 * res))]
    s = default_specs()
    s.plo[:] = [0.0, 0.0, 0.0]
    s.phi[:] = [2.0, 2.0, 7.0]
    s.ncells[:] = n
    for d in range(3):
        s.bclo[d] = s.bchi[d] = BC_SLIP
    s.bclo[2] = BC_NOSLIP
    s.order_scheme = 3
    s.num_redist = 1
    s.gravity[:] = [0.0, 0.0, 0.0]
    s.CFL = 0.3
    s.stress_update_scheme = USL
    cx, cy, ztop = 1.0, 1.0, 6.0

    def inside(X, Y, Z):
        return (np.abs(X - cx) <= 0.5) & (np.abs(Y - cy) <= 0.5) & (Z <= ztop)

    # lattice of the cells the column overlaps only (the full size has 100.7 M
    # particles in a 470 M-point domain lattice)
    box = ((cx - 0.5, cx + 0.5), (cy - 0.5, cy + 0.5), (0.0, ztop))
    pos, vp = lattice(s, (ppc, ppc, ppc), inside, box=box)
    w = omega0 * np.sin(np.pi * pos[2] / 12.0)
    vel = np.zeros_like(pos)
    vel[0] = -w * (pos[1] - cy)
    vel[1] = w * (pos[0] - cx)
    r, idata = make_records(pos, vp, 1000.0, 0, E=1e6, nu=0.3, vel=vel)
    return s, pos, r, idata


def random_block(seed, n=(12, 10, 8), order=2, periodic=(0, 0, 0), model=0, ppc=2,
                 scheme=USL, frac=0.6, speed=0.3, bcs=None):
    """Small seeded scene for parity tests: a jittered block of particles
    with random velocities and (model 0/2) random stresses."""
    rng = np.random.default_rng(seed)
    s = default_specs()
    s.plo[:] = [0.0, 0.0, 0.0]
    s.phi[:] = [n[0] / 16, n[1] / 16, n[2] / 16]
    s.ncells[:] = list(n)
    s.periodic[:] = list(periodic)
    for d in range(3):
        if periodic[d]:
            s.bclo[d] = s.bchi[d] = BC_PERIODIC
        elif bcs is not None:
            s.bclo[d], s.bchi[d] = bcs[d]
    s.order_scheme = order
    s.num_redist = 1
    s.gravity[:] = [0.0, -9.81, 0.0]
    s.fixed_timestep = 0
    s.CFL = 0.2
    s.stress_update_scheme = scheme
    lo = [s.phi[d] * (0.5 - frac / 2) for d in range(3)]
    hi = [s.phi[d] * (0.5 + frac / 2) for d in range(3)]
    for d in range(3):
        if periodic[d]:
            lo[d], hi[d] = s.plo[d], s.phi[d]

    def inside(X, Y, Z):
        return ((X >= lo[0]) & (X <= hi[0]) & (Y >= lo[1]) & (Y <= hi[1]) & (Z >= lo[2]) &
                (Z <= hi[2]))

    pos, vp = lattice(s, (ppc, ppc, ppc), inside)
    dx = (s.phi[0] - s.plo[0]) / n[0]
    pos = pos + rng.uniform(-0.2, 0.2, size=pos.shape) * dx
    for d in range(3):
        pos[d] = np.clip(pos[d], s.plo[d] + 1e-9, s.phi[d] - 1e-9)
    vel = rng.uniform(-speed, speed, size=pos.shape)
    if model == 1:
        r, idata = make_records(pos, vp, 1000.0, 1, K=2e4, gamma=7.0, mu=1e-3, vel=vel)
        r[R_JACOBIAN] = rng.uniform(0.98, 1.02, size=pos.shape[1])
    else:
        r, idata = make_records(pos, vp, 1000.0, model, E=1e4, nu=0.3, vel=vel)
        F = np.eye(3)[:, :, None] + rng.uniform(-0.01, 0.01, size=(3, 3, pos.shape[1]))
        r[R_DEFGRAD:R_DEFGRAD + 9] = F.reshape(9, -1)
        r[19:25] = rng.uniform(-50, 50, size=(6, pos.shape[1]))
        r[13:19] = rng.uniform(-1e-3, 1e-3, size=(6, pos.shape[1]))
    return s, pos, r, idata
