"""Pins for the CPU oracle (closed forms, invariants, brute force).

Each test checks the oracle against something the mathematics fixes, chosen
so a plausible restatement mistake (wrong coefficient, sign, index, shape
type, transposition) fails it.
"""
import numpy as np
import pytest

import oracle
from exagoop_amd import scenes
from exagoop_amd.capi import (BC_NOSLIP, BC_PARTIALSLIP, BC_PERIODIC, BC_SLIP, N_FRC, N_MASS,
                              N_VEL, R_DEFGRAD, R_DENSITY, R_JACOBIAN, R_MASS, R_PRESSURE,
                              R_STRAIN, R_STRAINRATE, R_STRESS, R_VEL, R_VELP, R_VOLUME,
                              default_specs)


# ---------------------------------------------------------------- 1-D splines
def textbook_cubic(z):
    a = abs(z)
    if a < 1:
        return 2.0 / 3.0 - z * z + 0.5 * a ** 3
    if a < 2:
        return (2 - a) ** 3 / 6.0
    return 0.0


def textbook_quad(z):
    a = abs(z)
    if a < 0.5:
        return 0.75 - z * z
    if a < 1.5:
        return 0.5 * (1.5 - a) ** 2
    return 0.0


@pytest.mark.parametrize("z", np.linspace(-1.99, 1.99, 41))
def test_interior_cubic_matches_textbook(z):
    # Truncated constants (constants.H:10,16,18) bound agreement at ~1e-10.
    assert abs(oracle.spline("cubic", 3, z) - textbook_cubic(z)) < 1e-10


@pytest.mark.parametrize("z", np.linspace(-1.49, 1.49, 41))
def test_interior_quadratic_matches_textbook(z):
    assert abs(oracle.spline("quad", 3, z) - textbook_quad(z)) < 1e-15


@pytest.mark.parametrize("kind,t", [(k, t) for k in ("quad", "cubic") for t in (1, 2, 3, 4)])
def test_spline_derivative_matches_finite_difference(kind, t):
    h = 1e-6
    support = 1.5 if kind == "quad" else 2.0
    for z in np.linspace(-support + 0.013, support - 0.013, 97):
        # skip points within h of a segment join
        joins = np.arange(-2, 2.5, 0.5)
        if np.min(np.abs(z - joins)) < 1e-3:
            continue
        fd = (oracle.spline(kind, t, z + h) - oracle.spline(kind, t, z - h)) / (2 * h)
        an = oracle.spline(kind, t, z, der=True)
        assert abs(fd - an) < 1e-5, (kind, t, z, fd, an)


def _one_d_sums(order, n, cell, x, dx=0.1):
    lo, hi = [0, 0, 0], [n - 1, n - 1, n - 1]
    if cell == 0:
        b0, b1 = 0, 3
    elif cell == n - 1:
        b0, b1 = -1, 2
    else:
        b0, b1 = -1, 3
    s = sd = sx = 0.0
    for l in range(b0, b1):
        N, g = oracle.basisval_and_grad([l, 0, 0], [cell, 5, 5], [x, 0.55, 0.55], [0, 0, 0],
                                        [dx] * 3, [order, 1, 1], [0, 1, 1], lo, hi)
        s += N / 0.25
        sd += g[0] / 0.25
        sx += N / 0.25 * (cell + l) * dx
    return s, sd, sx


@pytest.mark.parametrize("order", [2, 3])
def test_partition_of_unity_all_cells(order):
    n, dx = 10, 0.1
    for cell in range(n):
        for f in np.linspace(0.001, 0.999, 29):
            x = (cell + f) * dx
            s, sd, sx = _one_d_sums(order, n, cell, x, dx)
            assert abs(s - 1.0) < 1e-10
            assert abs(sd) < 1e-12
            if order == 3 or 2 <= cell <= n - 3:
                # cubic boundary splines reproduce linears; quadratic interior only
                assert abs(sx - x) < 1e-10


def test_quadratic_boundary_types_do_not_reproduce_linears():
    # Property of the reference's quadratic boundary tables (interpolants.H:
    # 121-146): constants are reproduced, linears are not in cells 0,1,n-2,n-1.
    n, dx = 10, 0.1
    _, _, sx = _one_d_sums(2, n, 0, 0.25 * dx, dx)
    assert abs(sx - 0.25 * dx) > 1e-4


def test_linear_basis():
    N, g = oracle.basisval_and_grad([1, 0, 1], [3, 4, 5], [0.325, 0.41, 0.58], [0, 0, 0],
                                    [0.1] * 3, [1, 1, 1], [0, 0, 0], [0] * 3, [9] * 3)
    r = np.array([0.25, 0.1, 0.8])
    assert np.isclose(N, r[0] * (1 - r[1]) * r[2], rtol=1e-14)
    assert np.allclose(g, [10 * (1 - r[1]) * r[2], r[0] * -10 * r[2], r[0] * (1 - r[1]) * 10],
                       rtol=1e-13)


# ------------------------------------------------------------------ applybc
def test_applybc_closed_forms():
    n = [0.0, 0.0, 1.0]
    v = [0.3, -0.4, -2.0]
    rc, out = oracle.applybc(v, 0.1, n, BC_NOSLIP)
    assert rc == 1 and np.all(out == 0)
    rc, out = oracle.applybc(v, 0.1, n, BC_SLIP)
    assert rc == 1 and np.allclose(out, [0.3, -0.4, 0.0])
    # partial slip, sliding: |vt| = 0.5 > mu*|vn| = 0.2 -> (|vt| + mu vn) t_hat
    rc, out = oracle.applybc(v, 0.1, n, BC_PARTIALSLIP)
    assert rc == 1 and np.allclose(out, np.array([0.3, -0.4, 0]) * (0.5 - 0.2) / 0.5)
    # partial slip, sticking
    rc, out = oracle.applybc(v, 0.5, n, BC_PARTIALSLIP)
    assert rc == 1 and np.all(out == 0)
    # partial slip, separating (vn > 0): untouched
    rc, out = oracle.applybc([0.3, -0.4, 2.0], 0.5, n, BC_PARTIALSLIP, preload=[7, 7, 7])
    assert rc == 0 and np.all(out == 7)
    rc, out = oracle.applybc(v, 0.5, n, BC_PERIODIC, preload=v)
    assert rc == 0 and np.allclose(out, v)
    rc, _ = oracle.applybc(v, 0.5, n, 4)
    assert rc == -1  # outflow aborts in the reference


# -------------------------------------------------------------- P2G by hand
def _single_particle_specs(order=2):
    s = default_specs()
    s.plo[:] = [0, 0, 0]
    s.phi[:] = [1.0, 1.0, 1.0]
    s.ncells[:] = [10, 10, 10]
    s.order_scheme = order
    s.num_redist = 1
    s.gravity[:] = [0.0, -9.81, 0.0]
    return s


def _one_particle(x, v, sigma, m=2.0, V=0.001):
    pos = np.array(x, dtype=float).reshape(3, 1)
    r, idata = scenes.make_records(pos, V, m / V, 0, E=1e4, nu=0.3)
    r[R_MASS] = m
    r[R_VOLUME] = V
    r[R_VEL:R_VEL + 3, 0] = v
    r[R_STRESS:R_STRESS + 6, 0] = sigma
    return pos, r, idata


def test_p2g_single_particle_by_hand():
    s = _single_particle_specs(order=2)
    x = [0.437, 0.512, 0.561]
    v = [0.3, -0.2, 0.1]
    sig = [10.0, 2.0, -3.0, 5.0, 1.5, -4.0]  # XX XY XZ YY YZ ZZ
    pos, r, idata = _one_particle(x, v, sig)
    o = oracle.Oracle(s)
    o.upload(pos, r, idata)
    o.fill_neighbors()
    o.p2g(1, 1, 1)
    m = o.nodal(N_MASS).reshape(11, 11, 11)  # [k, j, i]
    vel = [o.nodal(N_VEL + d).reshape(11, 11, 11) for d in range(3)]
    f = [o.nodal(N_FRC + d).reshape(11, 11, 11) for d in range(3)]
    S = np.array([[10, 2, -3], [2, 5, 1.5], [-3, 1.5, -4.0]])
    dx = 0.1
    iv = [int(np.floor(xx / dx)) for xx in x]
    tot = 0.0
    for K in range(iv[2] - 1, iv[2] + 3):
        for J in range(iv[1] - 1, iv[1] + 3):
            for I in range(iv[0] - 1, iv[0] + 3):
                z = [(x[0] - I * dx) / dx, (x[1] - J * dx) / dx, (x[2] - K * dx) / dx]
                w = [textbook_quad(zz) for zz in z]
                # textbook quadratic derivative
                def dq(zz):
                    a = abs(zz)
                    return -2 * zz if a < 0.5 else (-(1.5 - a) * np.sign(zz) if a < 1.5 else 0.0)
                dw = [dq(zz) / dx for zz in z]
                N = w[0] * w[1] * w[2]
                gN = np.array([dw[0] * w[1] * w[2], w[0] * dw[1] * w[2], w[0] * w[1] * dw[2]])
                assert abs(m[K, J, I] - 2.0 * N) < 1e-14
                tot += m[K, J, I]
                if N > 1e-12:
                    for d in range(3):
                        assert abs(vel[d][K, J, I] - v[d]) < 1e-13
                fexp = 2.0 * np.array([0, -9.81, 0]) * N - 0.001 * S @ gN
                for d in range(3):
                    assert abs(f[d][K, J, I] - fexp[d]) < 1e-12
    assert abs(tot - 2.0) < 1e-13


@pytest.mark.parametrize("order", [1, 2, 3])
def test_p2g_conservation_invariants(order):
    s, pos, r, idata = scenes.random_block(3, order=order)
    o = oracle.Oracle(s)
    o.upload(pos, r, idata)
    o.fill_neighbors()
    o.reset_nodal()
    o.p2g(1, 1, 1)
    m = o.nodal(N_MASS)
    v = np.stack([o.nodal(N_VEL + d) for d in range(3)])
    f = np.stack([o.nodal(N_FRC + d) for d in range(3)])
    M = r[R_MASS].sum()
    # cubic partition of unity holds to ~7e-11 per direction (truncated
    # constants.H literals), so the 3-D sum to ~2e-10
    tol = 3e-10 if order == 3 else 1e-11
    assert abs(m.sum() - M) / M < tol
    P = (r[R_MASS] * r[R_VEL:R_VEL + 3]).sum(axis=1)
    scale = (r[R_MASS] * np.abs(r[R_VEL:R_VEL + 3])).sum()
    assert np.allclose((m * v).sum(axis=1), P, rtol=0, atol=tol * scale)
    # internal forces sum to zero (sum grad N = 0); gravity sums to M g
    assert np.allclose(f.sum(axis=1), M * np.array([0, -9.81, 0]), rtol=1e-9, atol=1e-9)


def test_p2g_periodic_images_conserve_mass():
    s, pos, r, idata = scenes.random_block(5, order=2, periodic=(1, 0, 1))
    o = oracle.Oracle(s)
    o.upload(pos, r, idata)
    o.fill_neighbors()
    o.p2g(1, 1, 1)
    m = o.nodal(N_MASS).reshape(s.ncells[2] + 1, s.ncells[1] + 1, s.ncells[0] + 1)
    # periodic copies of the same node hold equal sums
    assert np.allclose(m[:, :, 0], m[:, :, -1], rtol=1e-12, atol=1e-15)
    assert np.allclose(m[0], m[-1], rtol=1e-12, atol=1e-15)
    # mass counted once on the unique nodes
    M = r[R_MASS].sum()
    assert abs(m[:-1, :, :-1].sum() - M) / M < 1e-10


# -------------------------------------------------------------- G2P gathers
@pytest.mark.parametrize("order", [2, 3])
def test_g2p_matches_textbook_gather(order):
    s, pos, r, idata = scenes.random_block(11, n=(12, 12, 12), order=order, frac=0.5)
    s.gravity[:] = [0, 0, 0]
    o = oracle.Oracle(s)
    o.upload(pos, r, idata)
    o.fill_neighbors()
    o.reset_nodal()
    o.p2g(1, 1, 1)
    o.backup_velocity()
    o.apply_nodal_bcs(0.0, 0.0)  # dv = v - v = 0
    vn = np.stack([o.nodal(N_VEL + d).reshape(13, 13, 13) for d in range(3)])
    o.g2p(1, 1, 0.0)
    _, rr, _, ids = o.download()
    dx = s.phi[0] / 12
    tb = textbook_quad if order == 2 else textbook_cubic
    h = 1e-6
    for q in range(0, pos.shape[1], 37):
        x = pos[:, q]
        iv = np.floor(x / dx).astype(int)
        if np.any(iv < 2) or np.any(iv > 9):
            continue  # interior cells only: textbook weights apply
        vexp = np.zeros(3)
        L = np.zeros((3, 3))
        for K in range(iv[2] - 1, iv[2] + 3):
            for J in range(iv[1] - 1, iv[1] + 3):
                for I in range(iv[0] - 1, iv[0] + 3):
                    z = (x - np.array([I, J, K]) * dx) / dx
                    w = [tb(zz) for zz in z]
                    dw = [(tb(zz + h) - tb(zz - h)) / (2 * h) / dx for zz in z]
                    N = w[0] * w[1] * w[2]
                    gN = np.array([dw[0] * w[1] * w[2], w[0] * dw[1] * w[2], w[0] * w[1] * dw[2]])
                    vI = vn[:, K, J, I]
                    vexp += N * vI
                    L += np.outer(vI, gN)
        assert np.allclose(rr[R_VELP:R_VELP + 3, q], vexp, rtol=1e-9, atol=1e-11)
        D = 0.5 * (L + L.T)
        Dp = [D[0, 0], D[0, 1], D[0, 2], D[1, 1], D[1, 2], D[2, 2]]
        assert np.allclose(rr[R_STRAINRATE:R_STRAINRATE + 6, q], Dp, rtol=1e-5, atol=1e-7)


# ----------------------------------------------------------- constitutive
def _stress_of(model, F=None, D=None, eps=None, J=1.0, E=1e4, nu=0.3, K=2e4, gamma=7.0, mu=1e-3,
               dt=0.0):
    s = _single_particle_specs(order=1)
    pos = np.array([[0.55], [0.55], [0.55]])
    r, idata = scenes.make_records(pos, 1e-3, 1000.0, model, E=E, nu=nu, K=K, gamma=gamma, mu=mu)
    if F is not None:
        r[R_DEFGRAD:R_DEFGRAD + 9, 0] = np.asarray(F).ravel()
    if D is not None:
        r[R_STRAINRATE:R_STRAINRATE + 6, 0] = D
    if eps is not None:
        r[R_STRAIN:R_STRAIN + 6, 0] = eps
    r[R_JACOBIAN, 0] = J
    o = oracle.Oracle(s)
    o.upload(pos, r, idata)
    o.stress(dt, 0.0)
    _, rr, _, _ = o.download()
    return rr[:, 0]


def test_hooke_uniaxial_strain():
    E, nu, e = 1e4, 0.3, 1e-3
    out = _stress_of(0, eps=[e, 0, 0, 0, 0, 0], E=E, nu=nu)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    G = E / (2 * (1 + nu))
    sig = out[R_STRESS:R_STRESS + 6]
    assert np.allclose(sig, [(lam + 2 * G) * e, 0, 0, lam * e, 0, lam * e], rtol=1e-13)
    out = _stress_of(0, eps=[0, 2e-3, 0, 0, 0, 0], E=E, nu=nu)  # tensor shear strain
    assert np.isclose(out[R_STRESS + 1], 2 * G * 2e-3, rtol=1e-13)


def test_strain_integration_and_applied_rate():
    D = [1.0, 0.5, -0.25, 2.0, 0.0, -1.0]
    out = _stress_of(0, D=D, dt=1e-3)
    assert np.allclose(out[R_STRAIN:R_STRAIN + 6], np.array(D) * 1e-3, rtol=1e-15)


def test_tait_newtonian():
    J, K, g, mu = 0.99, 2e4, 7.0, 1e-3
    D = [0.1, 0.2, 0.3, -0.4, 0.5, 0.6]
    out = _stress_of(1, D=D, J=J, K=K, gamma=g, mu=mu)
    p = K * ((1 / J) ** g - 1)  # no 1/gamma factor (mpm_particle_container.cpp:115-120)
    assert np.isclose(out[R_PRESSURE], p, rtol=1e-14)
    tr = D[0] + D[3] + D[5]
    exp = [2 * mu * (D[0] - tr / 3) - p, 2 * mu * D[1], 2 * mu * D[2], 2 * mu * (D[3] - tr / 3) - p,
           2 * mu * D[4], 2 * mu * (D[5] - tr / 3) - p]
    assert np.allclose(out[R_STRESS:R_STRESS + 6], exp, rtol=1e-13)


def test_neo_hookean_uniaxial_stretch():
    E, nu, lam_s = 1e4, 0.3, 1.1
    out = _stress_of(2, F=np.diag([lam_s, 1, 1]), E=E, nu=nu)
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    vol = lam * np.log(lam_s) / lam_s
    exp = [mu / lam_s * (lam_s ** 2 - 1) + vol, 0, 0, vol, 0, vol]
    assert np.allclose(out[R_STRESS:R_STRESS + 6], exp, rtol=1e-13, atol=1e-12)
    out = _stress_of(2, F=np.eye(3))
    assert np.allclose(out[R_STRESS:R_STRESS + 6], 0, atol=1e-12)


def test_volume_update():
    s = _single_particle_specs(order=1)
    pos = np.array([[0.55], [0.55], [0.55]])
    r, idata = scenes.make_records(pos, 1e-3, 1000.0, 0, E=1e4, nu=0.3)
    F = np.array([[1.1, 0.2, 0.0], [0.0, 0.9, 0.1], [0.05, 0.0, 1.0]])
    r[R_DEFGRAD:R_DEFGRAD + 9, 0] = F.ravel()
    o = oracle.Oracle(s)
    o.upload(pos, r, idata)
    o.update_volume()
    _, rr, _, _ = o.download()
    J = np.linalg.det(F)
    assert np.isclose(rr[R_JACOBIAN, 0], J, rtol=1e-14)
    assert np.isclose(rr[R_VOLUME, 0], r[R_VOLUME, 0] * J, rtol=1e-14)
    assert np.isclose(rr[R_DENSITY, 0], r[R_MASS, 0] / (r[R_VOLUME, 0] * J), rtol=1e-14)


def test_dt_closed_form():
    s, pos, r, idata = scenes.dam_break(nx=16, ny=8, nz=8)
    r[R_VEL, 0] = 3.0
    o = oracle.Oracle(s)
    o.upload(pos, r, idata)
    c = np.sqrt(2e6 / 1000.0)
    assert np.isclose(o.calculate_dt(), 0.3 * (2.0 / 16) / (c + 3.0), rtol=1e-15)


# --------------------------------------------------------------- whole step
def test_periodic_momentum_conservation():
    """Fully periodic, no gravity: PIC/FLIP USL conserves total momentum."""
    s, pos, r, idata = scenes.random_block(7, n=(8, 8, 8), order=2, periodic=(1, 1, 1))
    s.gravity[:] = [0, 0, 0]
    o = oracle.Oracle(s)
    o.upload(pos, r, idata)
    o.init_internal_forces()
    P0 = o.energies()[2:5]
    o.step(5)
    P1 = o.energies()[2:5]
    assert np.allclose(P1, P0, rtol=1e-11, atol=1e-13)


def test_nodal_bc_corner_quirk():
    """nodal_bcs applies each direction to the accumulated relative velocity
    and the last direction's write wins (nodal_data_ops.cpp:720-763)."""
    s = _single_particle_specs(order=1)
    s.gravity[:] = [0, 0, 0]
    s.bclo[0] = BC_SLIP
    s.bclo[1] = BC_NOSLIP
    s.wall_vel_lo[0:3] = [0.0, 0.5, 0.0]  # x-lo wall moves in +y
    pos, r, idata = _one_particle([0.03, 0.03, 0.55], [1.0, 2.0, 3.0], [0] * 6)
    o = oracle.Oracle(s)
    o.upload(pos, r, idata)
    o.fill_neighbors()
    o.reset_nodal()
    o.p2g(1, 1, 0)
    o.apply_nodal_bcs(0.0, 0.0)
    v = np.stack([o.nodal(N_VEL + d).reshape(11, 11, 11) for d in range(3)])
    # corner node (0,0,k): x-lo subtracts w, y-lo (no-slip) zeroes the
    # relative velocity, result = 0 + w_y-lo = 0
    assert np.allclose(v[:, 5, 0, 0], 0.0)
    # x-lo face node (0,1,5): slip on v - w: vx removed, then + w
    assert np.allclose(v[:, 5, 1, 0], [0.0, 2.0, 3.0])
