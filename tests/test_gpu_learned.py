"""GPU checks of the learned coarse-step correction through the C ABI: the
BF16 MFMA network's forward pass, gradient and Adam step against the CPU
oracle (oracle/oracle_mlp.cpp, pinned in test_mlp_oracle.py), the nodal
features against the oracle's gather, the learned step's plumbing (one
uncorrected coarse step of one fine dt is the plain step; a rejected
correction falls back to the plain steps exactly) and training."""
import numpy as np
import pytest

import oracle
from exagoop_amd import capi, scenes

pytestmark = pytest.mark.gpu

NP = capi.MLP_NPARAMS


def say(msg):
    print(msg, flush=True)


def rand_params(seed, scale=0.15):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(NP) * scale).astype(np.float32)


def rand_rows(seed, n):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, capi.MLP_IN), dtype=np.float32)
    X[:, :108] = rng.standard_normal((n, 108)).astype(np.float32)
    T = rng.standard_normal((n, 3)).astype(np.float32)
    return X, T


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300)


def test_forward_matches_oracle(gpu):
    say("[learned] forward vs oracle")
    p = rand_params(1)
    X, _ = rand_rows(2, 1000)  # not a multiple of the 32-row tile
    net = capi.MLP(p)
    y = net.forward(X)
    ref = oracle.mlp_forward(p, X, exact=False)
    # same BF16 rounding points; FP32 sums in a different order
    assert rel(y, ref) <= 2e-3, rel(y, ref)
    assert rel(y, oracle.mlp_forward(p, X, exact=True)) <= 5e-2


def test_gradient_matches_oracle(gpu):
    say("[learned] gradient vs oracle")
    p = rand_params(3)
    X, T = rand_rows(4, 777)
    net = capi.MLP(p)
    loss, g = net.gradient(X, T)
    lref, gref = oracle.mlp_grad(p, X, T, exact=False)
    assert abs(loss - lref) <= 2e-3 * abs(lref), (loss, lref)
    assert rel(g, gref) <= 1e-2, rel(g, gref)
    # every parameter block carries gradient
    for lo, hi in ((0, 6912), (6912, 6976), (6976, 11072), (11072, 11136), (11136, 11328),
                   (11328, 11331)):
        assert rel(g[lo:hi], gref[lo:hi]) <= 2e-2, (lo, hi, rel(g[lo:hi], gref[lo:hi]))


def test_gradient_is_reproducible(gpu):
    """The gradient is summed without atomics in a fixed order (per-wavefront
    partial rows, then rows in order): the same bits on every call."""
    say("[learned] gradient twice, bitwise")
    p = rand_params(11)
    X, T = rand_rows(12, 3000)
    net = capi.MLP(p)
    l1, g1 = net.gradient(X, T)
    l2, g2 = net.gradient(X, T)
    assert l1 == l2
    assert np.array_equal(g1, g2)


def test_gradient_of_a_batch_above_one_launch(gpu):
    """20 000 rows = 625 tiles: the first 512 tiles store their partial rows,
    a second launch adds the other 113 into rows 0..112."""
    say("[learned] gradient, 20000 rows (two launches)")
    p = rand_params(13)
    X, T = rand_rows(14, 20000)
    net = capi.MLP(p)
    loss, g = net.gradient(X, T)
    lref, gref = oracle.mlp_grad(p, X, T, exact=False)
    assert abs(loss - lref) <= 2e-3 * abs(lref), (loss, lref)
    assert rel(g, gref) <= 1e-2, rel(g, gref)
    _, g2 = net.gradient(X, T)
    assert np.array_equal(g, g2)


def test_adam_matches_oracle(gpu):
    say("[learned] adam vs oracle")
    p = rand_params(5)
    net = capi.MLP(p)
    m = np.zeros(NP, np.float32)
    v = np.zeros(NP, np.float32)
    pr = p.copy()
    rng = np.random.default_rng(6)
    for t in range(1, 4):
        g = rng.standard_normal(NP).astype(np.float32)
        net.adam(g, 1e-3, t)
        oracle.mlp_adam(pr, m, v, g, 1e-3, t)
    got, _ = net.get_params()
    assert np.max(np.abs(got - pr)) <= 1e-6


def small_fluid(seed=7):
    s, pos, r, idata = scenes.dam_break(nx=32, ny=16, nz=16)
    return s, pos, r, idata


def test_features_match_oracle(gpu):
    say("[learned] node features vs oracle gather")
    s, pos, r, idata = small_fluid()
    g = capi.Context(s)
    g.upload(pos, r, idata)
    g.init_internal_forces()
    g.step(3)
    mass = g.nodal(capi.N_MASS)
    vel = np.stack([g.nodal(capi.N_VEL + d) for d in range(3)])
    nodes = np.flatnonzero(mass >= s.mass_tolerance)
    nodes = np.concatenate([nodes[:50], nodes[-50:], [0, g.nn - 1]])
    net = capi.MLP(rand_params(8), scales=(2.0, 0.5, 1.0))
    X = capi.learn_features(g, net, nodes)
    ref = oracle.mlp_features(mass, vel, list(s.ncells), nodes, 2.0, 0.5)
    assert np.array_equal(X, oracle.bf16_round(ref))


@pytest.mark.parametrize("blend", [0, 1])
def test_uncorrected_single_fine_coarse_step_is_the_plain_step(gpu, blend):
    # with nfine = 1 the coarse blend alpha^nfine is alpha itself
    say(f"[learned] coarse step of one fine dt, no network == plain step (blend option {blend})")
    s, pos, r, idata = small_fluid()
    a = capi.Context(s)
    b = capi.Context(s)
    for c in (a, b):
        c.upload(pos, r, idata)
        c.init_internal_forces()
        c.step(2)
    a.set_option(capi.OPT_COARSE_BLEND, blend)
    fb = capi.step_learned(a, None, 5, 1)
    b.step(5)
    assert fb == 0
    pa, ra, _, ia = a.download()
    pb, rb, _, ib = b.download()
    assert np.array_equal(ia, ib)
    assert rel(pa, pb) <= 1e-13
    assert rel(ra, rb) <= 1e-12
    assert a.time()[:2] == b.time()[:2] or abs(a.time()[0] - b.time()[0]) <= 1e-15


def test_accepted_zero_correction_of_one_fine_dt_is_the_plain_step(gpu):
    say("[learned] zero network, coarse step of one fine dt, accepted == plain step")
    s, pos, r, idata = small_fluid()
    a = capi.Context(s)
    b = capi.Context(s)
    for c in (a, b):
        c.upload(pos, r, idata)
        c.init_internal_forces()
        c.step(2)
    net = capi.MLP(np.zeros(NP, np.float32), scales=(1.0, 1.0, 1.0))  # correction 0
    fb = capi.step_learned(a, net, 5, 1)
    b.step(5)
    assert fb == 0
    pa, ra, _, ia = a.download()
    pb, rb, _, ib = b.download()
    assert np.array_equal(ia, ib)
    assert rel(pa, pb) <= 1e-13
    assert rel(ra, rb) <= 1e-12
    # the accepted steps' time advances (deferred into the next dt launch)
    assert a.time()[1] == b.time()[1]
    assert abs(a.time()[0] - b.time()[0]) <= 1e-15


def test_rejected_correction_falls_back_to_plain_steps(gpu):
    say("[learned] guard rejects a wild network -> plain steps")
    s, pos, r, idata = small_fluid()
    a = capi.Context(s)
    b = capi.Context(s)
    for c in (a, b):
        c.upload(pos, r, idata)
        c.init_internal_forces()
        c.step(2)
    p = np.zeros(NP, np.float32)
    p[-3:] = [50.0, -40.0, 30.0]  # b3: a large constant correction
    net = capi.MLP(p, scales=(1.0, 1.0, 1.0))
    fb = capi.step_learned(a, net, 3, 4)
    b.step(12)
    assert fb == 3
    pa, ra, _, ia = a.download()
    pb, rb, _, ib = b.download()
    assert np.array_equal(ia, ib)
    assert rel(pa, pb) <= 1e-13
    assert rel(ra, rb) <= 1e-12


def test_training_fits_a_smooth_map(gpu):
    say("[learned] training on a synthetic smooth map")
    rng = np.random.default_rng(9)
    n = 8192
    X = np.zeros((n, capi.MLP_IN), dtype=np.float32)
    X[:, :108] = rng.uniform(-1, 1, (n, 108)).astype(np.float32)
    A = rng.standard_normal((108, 3)) / np.sqrt(108)
    T = (np.tanh(X[:, :108] @ A) * 0.02).astype(np.float32)  # targets in m/s, dv_ref 0.02
    ds = capi.Dataset(n)
    ds.add(X, T)
    assert ds.size() == n
    assert abs(ds.target_rms() - np.sqrt(np.mean(T.astype(np.float64) ** 2))) <= 1e-6
    net = capi.MLP(rand_params(10, 0.05), scales=(1.0, 1.0, 0.02))
    losses = net.train(ds, 12, 256, 2e-3, seed=1)
    assert np.all(np.isfinite(losses))
    assert losses[-1] <= 0.2 * losses[0], losses
    y = net.forward(X[:512]) * 0.02
    assert rel(y, T[:512]) <= 0.35


def test_collect_rows_from_the_solver(gpu):
    say("[learned] training rows from the plain solver")
    s, pos, r, idata = small_fluid()
    g = capi.Context(s)
    g.upload(pos, r, idata)
    g.init_internal_forces()
    g.step(10)
    net = capi.MLP(np.zeros(NP, np.float32), scales=(1.0, 1.0, 1.0))
    ds = capi.Dataset(20000)
    t0 = g.time()
    capi.learn_collect(g, net, ds, 4, 3000, 11)
    t1 = g.time()
    assert t1[1] == t0[1] + 4  # the context advanced four plain steps
    n = ds.size()
    assert 1000 <= n <= 3000, n
    assert np.isfinite(ds.target_rms()) and ds.target_rms() > 0


def test_learned_step_applies_features_then_forward(gpu):
    """The learned step gathers each active node's features inside its
    inference kernel.  After one accepted coarse step the nodal velocity
    minus that of the same coarse step without a network must be
    dv_ref * forward(features) at the active nodes (features from the
    separate feature kernel, forward from the separate inference kernel) and
    zero elsewhere.  The two runs deposit with FP64 atomics, so their nodal
    states agree to rounding, not bitwise."""
    from exagoop_amd import learned
    say("[learned] fused features + inference + correction vs the separate kernels")
    net = learned.load_net()
    _, sc = net.get_params()
    s, _, _, _ = learned.scene(1005)
    runs = {}
    for name, model in (("learned", net), ("coarse", None)):
        g = learned.context(1005)
        g.step(40)
        fb = capi.step_learned(g, model, 1, learned.NFINE)
        assert fb == 0
        runs[name] = g
    a, b = runs["learned"], runs["coarse"]
    mass = b.nodal(capi.N_MASS)
    nodes = np.flatnonzero(mass >= s.mass_tolerance)
    assert nodes.size > 1000
    Y = net.forward(capi.learn_features(b, net, nodes))
    dv = Y.astype(np.float64) * sc[2]
    va = np.stack([a.nodal(capi.N_VEL + d) for d in range(3)])
    vb = np.stack([b.nodal(capi.N_VEL + d) for d in range(3)])
    diff = (va - vb)[:, nodes].T
    assert np.abs(dv).max() > 0
    assert rel(diff, dv) <= 1e-3, rel(diff, dv)
    rest = np.ones(mass.size, bool)
    rest[nodes] = False
    assert np.abs((va - vb)[:, rest]).max() <= 1e-9 * max(np.abs(vb).max(), 1.0)
    for g in (a, b):
        g.close()


def wild_net():
    p = np.zeros(NP, np.float32)
    p[-3:] = [50.0, -40.0, 30.0]  # b3: a large constant correction, always rejected
    return capi.MLP(p, scales=(1.0, 1.0, 1.0))


@pytest.mark.parametrize("case", ["trained_accepts", "wild_rejects", "tight_guard"])
def test_speculative_steps_equal_one_step_calls(gpu, case):
    """Within one call the learned step enqueues step k + 1 before reading
    step k's decision and undoes it when step k is rejected.  One call of 8
    steps must give the bits of 8 calls of one step (no speculation across
    calls), with the same fallbacks; the deterministic P2G makes the runs
    bitwise comparable."""
    from exagoop_amd import learned
    say(f"[learned] speculative steps vs one-step calls: {case}")
    net = {"trained_accepts": learned.load_net, "wild_rejects": wild_net,
           "tight_guard": learned.load_net}[case]()
    guards = (1e-3, 0.02) if case == "tight_guard" else (0.05, 0.5)
    runs = []
    for calls in ((8,), (1,) * 8):
        g = learned.context(1007)
        g.set_option(capi.OPT_DETERMINISTIC, 1)
        g.step(20)
        fb = sum(capi.step_learned(g, net, n, learned.NFINE, guards) for n in calls)
        p, r, i, ids = g.download()
        runs.append((fb, p, r, ids, g.time()))
        g.close()
    (fa, pa, ra, ida, ta), (fb, pb, rb, idb, tb) = runs
    say(f"    fallbacks {fa} (one call) {fb} (one-step calls)")
    if case == "trained_accepts":
        assert fa == 0
    if case == "wild_rejects":
        assert fa == 8
    assert fa == fb
    assert np.array_equal(ida, idb)
    assert pa.tobytes() == pb.tobytes()
    assert ra.tobytes() == rb.tobytes()
    assert ta[:2] == tb[:2]


@pytest.mark.parametrize("nsteps", [100, 800])
def test_trained_network_gains_over_the_uncorrected_coarse_step(gpu, nsteps):
    """The committed network (exagoop_amd/data/learned_dam_break.npz, trained
    by exagoop_amd.learned on seeds 0..63 over their first 800 fine steps) on
    a held-out dam-break geometry: nsteps fine steps of dt = 2.5e-5 against
    nsteps / 4 coarse steps of 1e-4 with the correction, and without it both
    with the plain PIC/FLIP blend and with the blend alpha^4 of the four fine
    steps (MPM_OPT_COARSE_BLEND), the better of the two uncorrected forms."""
    from exagoop_amd import learned
    say(f"[learned] trained network vs uncorrected coarse steps, held-out seed 1003, "
        f"{nsteps} fine steps")
    net = learned.load_net()
    seed = 1003
    g = learned.context(seed)
    g.step(nsteps)
    vf, idf = learned.velocities(g)
    g.close()
    errs = {}
    for name, model, blend in (("learned", net, 0), ("coarse", None, 0),
                               ("coarse_blend", None, 1)):
        g = learned.context(seed)
        g.set_option(capi.OPT_COARSE_BLEND, blend)
        fb = capi.step_learned(g, model, nsteps // learned.NFINE, learned.NFINE)
        v, ids = learned.velocities(g)
        g.close()
        assert np.array_equal(ids, idf)
        assert fb == 0
        errs[name] = learned.rel_l2(v, vf)
    say(f"    errors {errs}")
    assert errs["coarse_blend"] < errs["coarse"], errs
    assert errs["learned"] <= 0.5 * errs["coarse_blend"], errs
