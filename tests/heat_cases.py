"""The reference's heat-conduction validation cases (Tests/1D_Heat_Conduction,
Tests/1D_Heat_Conduction_Convective, Tests/1D_Heat_Conduction_HeatFlux,
Tests/2D_Heat_Conduction, Tests/2D_Heat_Conduction_Cylinder_Dirichlet:
PreProcess/config.json for the set-up,
PostProcess/validate.py for the time, the analytic solution and the RMS
tolerance), set up for the 3-D step, and their analytic solutions.

The 1-D and 2-D builds' domains become slabs one or a few cells thick along
the missing directions, with adiabatic temperature walls and slip walls
there: with fields uniform across the slab these walls carry no heat, so the
3-D solution is the 1-D / 2-D one.  The missing directions' cells are 1
long, so particle volumes and masses are the lower-dimensional builds'
lengths or areas (mass_tolerance then cuts at the same fraction of a
particle).  A slab is at least five cells thick where the order is
quadratic, so that direction keeps the quadratic spline (the small-grid
fallback of mpm_init.cpp:180-233 makes a thinner non-periodic direction
linear).  Material: rho = c_p = k = 1 (alpha = 1),
elastic E = 1e6, nu = 0.3, at rest without gravity; initial T = 0.
"""
import numpy as np

from exagoop_amd import capi, scenes

try:
    from scipy.integrate import quad
    from scipy.optimize import brentq
except ImportError:  # pragma: no cover - scipy is in the image
    quad = brentq = None


def dirichlet_1d(x, t, T0=0.0, T1=1.0, nterms=100):
    """T(0) = T0, T(1) = T1, T(x, 0) = T0 on [0, 1], alpha = 1."""
    s = np.zeros_like(x)
    for n in range(1, nterms + 1):
        s += (-1.0) ** n / (n * np.pi) * np.exp(-(n * np.pi) ** 2 * t) * np.sin(n * np.pi * x)
    return T0 + (T1 - T0) * x + 2.0 * (T1 - T0) * s


def convective_1d(x, t, T_wall=1.0, h=2.0, T_inf=0.0, k=1.0, L=1.0, nterms=60):
    """T(0) = T_wall, -k T'(L) = h (T(L) - T_inf), T(x, 0) = 0, alpha = 1."""
    b = h / k
    T_ss = lambda z: T_wall + (T_inf - T_wall) * b * z / (1.0 + b * L)  # noqa: E731
    T = T_ss(x)
    for n in range(1, nterms + 1):
        # lambda cos(lambda L) + b sin(lambda L) = 0, root in ((n - 1/2) pi, n pi) / L
        f = lambda lam: lam * np.cos(lam * L) + b * np.sin(lam * L)  # noqa: E731
        lam = brentq(f, (n - 0.5) * np.pi / L + 1e-12, n * np.pi / L - 1e-12, xtol=1e-15)
        num = quad(lambda z: -T_ss(z) * np.sin(lam * z), 0.0, L, limit=200)[0]
        den = quad(lambda z: np.sin(lam * z) ** 2, 0.0, L, limit=200)[0]
        T = T + num / den * np.sin(lam * x) * np.exp(-lam * lam * t)
    return T


def heatflux_1d(x, t, q=1.0, k=1.0, L=1.0, nterms=200):
    """T(0) = 0, k T'(L) = q (heat entering at x = L), T(x, 0) = 0."""
    T = q * x / k
    for n in range(1, nterms + 1):
        lam = (2 * n - 1) * np.pi / (2 * L)
        # C_n = -(2 / L) int_0^L (q z / k) sin(lam z) dz
        integral = (q / k) * (np.sin(lam * L) - lam * L * np.cos(lam * L)) / lam ** 2
        T = T + (-(2.0 / L) * integral) * np.sin(lam * x) * np.exp(-lam * lam * t)
    return T


def dirichlet_2d(x, y, t, T0=0.0, T1=1.0, L=1.0, H=1.0, nterms=25):
    """All four walls at T1, T(x, y, 0) = T0, alpha = 1."""
    T = np.full_like(x, T1)
    for i in range(1, 2 * nterms, 2):
        for j in range(1, 2 * nterms, 2):
            decay = np.exp(-np.pi ** 2 * (i * i / L ** 2 + j * j / H ** 2) * t)
            T += (T0 - T1) * 16.0 / np.pi ** 2 * decay / (i * j) * (
                np.sin(i * np.pi * x / L) * np.sin(j * np.pi * y / H))
    return T


def laplace_cylinder(x, y, R=0.15, cx=0.5, cy=0.5, T_in=1.0, T_out=0.0, N=400):
    """Steady T on [0, 1]^2 with T = T_in inside the circle of radius R and
    T = T_out on the square's edges: second-order finite differences on an
    N x N grid, bilinear interpolation to the points."""
    from scipy.interpolate import RegularGridInterpolator
    from scipy.sparse import lil_matrix
    from scipy.sparse.linalg import spsolve

    xs = np.linspace(0.0, 1.0, N + 1)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    inner = (X - cx) ** 2 + (Y - cy) ** 2 <= R * R
    edge = np.zeros_like(inner)
    edge[0, :] = edge[-1, :] = edge[:, 0] = edge[:, -1] = True
    idx = np.arange((N + 1) ** 2).reshape(N + 1, N + 1)
    A = lil_matrix(((N + 1) ** 2, (N + 1) ** 2))
    b = np.zeros((N + 1) ** 2)
    for i in range(N + 1):
        for j in range(N + 1):
            k = idx[i, j]
            if inner[i, j] or edge[i, j]:
                A[k, k] = 1.0
                b[k] = T_in if inner[i, j] else T_out
            else:
                A[k, k] = -4.0
                for a, c in ((i + 1, j), (i - 1, j), (i, j + 1), (i, j - 1)):
                    A[k, idx[a, c]] = 1.0
    T = spsolve(A.tocsr(), b).reshape(N + 1, N + 1)
    return RegularGridInterpolator((xs, xs), T)(np.stack([x, y], axis=1))


# name: (cells, extent, ppc, order, scheme, alpha, periodic, x/y temperature walls, steps,
#        dt, RMS tolerance; body: an isothermal cylinder along z)
CASES = {
    "1D_Heat_Conduction": dict(
        cells=(100, 1, 1), extent=(1.0, 1.0, 1.0), ppc=(2, 1, 1), order=1, scheme=capi.MUSL,
        alpha=1.0, periodic=(0, 0, 0), num_redist=10,
        walls={(0, 0): (capi.BC_TEMP_ISOTHERMAL, "T_wall", 0.0),
               (0, 1): (capi.BC_TEMP_ISOTHERMAL, "T_wall", 1.0)},
        steps=5000, dt=1e-5, tol=1e-2),
    "1D_Heat_Conduction_Convective": dict(
        cells=(40, 4, 5), extent=(1.0, 0.1, 5.0), ppc=(2, 2, 1), order=2, scheme=capi.MUSL,
        alpha=1.0, periodic=(0, 0, 0), num_redist=10,
        walls={(0, 0): (capi.BC_TEMP_ISOTHERMAL, "T_wall", 1.0),
               (0, 1): (capi.BC_TEMP_CONVECTION, "h", 2.0, "T_inf", 0.0)},
        steps=50000, dt=1e-5, tol=1e-2),
    "1D_Heat_Conduction_HeatFlux": dict(
        cells=(40, 4, 1), extent=(1.0, 0.1, 1.0), ppc=(2, 2, 1), order=1, scheme=capi.MUSL,
        alpha=1.0, periodic=(0, 0, 0), num_redist=10,
        walls={(0, 0): (capi.BC_TEMP_ISOTHERMAL, "T_wall", 0.0),
               (0, 1): (capi.BC_TEMP_USERFLUX, "flux", 1.0)},
        steps=50000, dt=1e-5, tol=1e-2),
    "2D_Heat_Conduction": dict(
        cells=(32, 32, 1), extent=(1.0, 1.0, 1.0), ppc=(2, 2, 1), order=1, scheme=capi.MUSL,
        alpha=1.0, periodic=(0, 1, 0), num_redist=1,
        walls={(0, 0): (capi.BC_TEMP_ISOTHERMAL, "T_wall", 1.0),
               (0, 1): (capi.BC_TEMP_ISOTHERMAL, "T_wall", 1.0),
               (1, 0): (capi.BC_TEMP_ISOTHERMAL, "T_wall", 1.0),
               (1, 1): (capi.BC_TEMP_ISOTHERMAL, "T_wall", 1.0)},
        steps=5000, dt=1e-5, tol=5e-2),
    "2D_Heat_Conduction_Cylinder_Dirichlet": dict(
        cells=(128, 128, 1), extent=(1.0, 1.0, 1.0), ppc=(2, 2, 1), order=1,
        scheme=capi.MUSL, alpha=1.0, periodic=(0, 0, 0), num_redist=10,
        walls={(0, 0): (capi.BC_TEMP_ISOTHERMAL, "T_wall", 0.0),
               (0, 1): (capi.BC_TEMP_ISOTHERMAL, "T_wall", 0.0),
               (1, 0): (capi.BC_TEMP_ISOTHERMAL, "T_wall", 0.0),
               (1, 1): (capi.BC_TEMP_ISOTHERMAL, "T_wall", 0.0)},
        body=dict(center=(0.5, 0.5), radius=0.15, T_wall=1.0),
        steps=10000, dt=1e-5, tol=1e-1),
}


def build(name):
    """specs, thermal specs, pos, rdata, idata, thermal fields of a case."""
    c = CASES[name]
    s = capi.default_specs()
    s.plo[:] = [0.0, 0.0, 0.0]
    s.phi[:] = list(c["extent"])
    s.ncells[:] = list(c["cells"])
    s.periodic[:] = list(c["periodic"])
    for d in range(3):
        s.bclo[d] = s.bchi[d] = capi.BC_PERIODIC if c["periodic"][d] else capi.BC_SLIP
    if "body" not in c:
        s.bclo[0] = s.bchi[0] = capi.BC_NOSLIP
    s.order_scheme = c["order"]
    s.stress_update_scheme = c["scheme"]
    s.alpha_pic_flip = c["alpha"]
    s.num_redist = c["num_redist"]
    s.gravity[:] = [0.0, 0.0, 0.0]
    s.fixed_timestep = 1
    s.timestep = c["dt"]
    s.CFL = 0.1
    ts = capi.default_thermal_specs(s)
    for (d, hi), spec in c["walls"].items():
        if hi:
            ts.bchi[d] = spec[0]
        else:
            ts.bclo[d] = spec[0]
        for name_, value in zip(spec[1::2], spec[2::2]):
            getattr(ts, name_ + ("_hi" if hi else "_lo"))[d] = value
    pos, vp = scenes.lattice(s, c["ppc"], lambda X, Y, Z: np.ones_like(X, bool))
    r, idata = scenes.make_records(pos, vp, 1.0, 0, E=1e6, nu=0.3)
    th = np.zeros((capi.NTHERMAL, pos.shape[1]))
    th[capi.TH_SPECIFIC_HEAT] = 1.0
    th[capi.TH_CONDUCTIVITY] = 1.0
    return s, ts, pos, r, idata, th


def exact(name, pos, t):
    if name == "2D_Heat_Conduction_Cylinder_Dirichlet":
        return laplace_cylinder(pos[0], pos[1])
    if name == "1D_Heat_Conduction":
        return dirichlet_1d(pos[0], t)
    if name == "1D_Heat_Conduction_Convective":
        return convective_1d(pos[0], t)
    if name == "1D_Heat_Conduction_HeatFlux":
        return heatflux_1d(pos[0], t)
    return dirichlet_2d(pos[0], pos[1], t)


def run(ctx_factory, name, steps=None):
    """Runs a case on a context (capi.Context or oracle.Oracle); returns the
    RMS error of the particle temperatures against the analytic solution,
    the context and the positions."""
    s, ts, pos, r, idata, th = build(name)
    ctx = ctx_factory(s)
    ctx.upload(pos, r, idata)
    ctx.enable_thermal(ts)
    ctx.upload_thermal(th)
    body = CASES[name].get("body")
    if body:
        # the level-set body (a "sphere" of the 2-D build: a cylinder along
        # z here), its particles removed as removeParticlesInsideEB does
        cx, cy = body["center"]
        phi = capi.levelset_analytic(s, capi.LS_CYLINDER, [cx, cy, 0.0, body["radius"], 0.0, 2])
        ctx.add_levelset_body(phi, capi.BC_SLIP, 0.0)
        ctx.set_levelset_body_temp(0, capi.LS_TEMP_ISOTHERMAL, T_wall=body["T_wall"])
        ctx.remove_particles_inside_levelset()
    ctx.init_internal_forces()
    ctx.step(CASES[name]["steps"] if steps is None else steps)
    T = ctx.download_thermal()[capi.TH_TEMPERATURE]
    p, _, _, ids = ctx.download()
    t = ctx.time()[0]
    Tp = T[ids]
    if body:
        # the validator's ring: distance from the axis in [R, 0.5]
        rr = np.hypot(p[0] - body["center"][0], p[1] - body["center"][1])
        keep = (rr >= body["radius"]) & (rr <= 0.5)
        p, Tp = p[:, keep], Tp[keep]
    err = Tp - exact(name, p, t)
    return float(np.sqrt(np.mean(err ** 2))), ctx
