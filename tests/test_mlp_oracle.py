"""Pins of the learned-correction oracle (oracle/oracle_mlp.cpp) on the CPU.

The network, loss and optimiser are checked against torch's own float64
layers, autograd and Adam (an independent implementation of the same
mathematics), the backward pass against central finite differences, the
bfloat16 rounding against torch's conversion, and the stencil features
against a brute-force gather on a small grid."""
import numpy as np
import pytest
import torch

import oracle

NP = oracle.MLP_NPARAMS


def rand_params(seed, scale=0.2):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(NP) * scale).astype(np.float32)
    return p


def unpack(p):
    p = torch.tensor(p, dtype=torch.float64)
    o = 0
    W1 = p[o:o + 108 * 64].reshape(108, 64); o += 108 * 64
    b1 = p[o:o + 64]; o += 64
    W2 = p[o:o + 64 * 64].reshape(64, 64); o += 64 * 64
    b2 = p[o:o + 64]; o += 64
    W3 = p[o:o + 64 * 3].reshape(64, 3); o += 64 * 3
    b3 = p[o:o + 3]; o += 3
    assert o == NP
    return W1, b1, W2, b2, W3, b3


def torch_forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = torch.tensor(X[:, :108], dtype=torch.float64)
    g = torch.nn.functional.gelu
    h1 = g(x @ W1 + b1, approximate="tanh")
    h2 = g(h1 @ W2 + b2, approximate="tanh")
    return h2 @ W3 + b3


def rand_rows(seed, n):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, 128), dtype=np.float32)
    X[:, :108] = rng.standard_normal((n, 108)).astype(np.float32)
    T = rng.standard_normal((n, 3)).astype(np.float32)
    return X, T


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-30, 30, 2000),
                        [0.0, -0.0, 1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 65504.0, 1e-40, np.inf,
                         -np.inf]]).astype(np.float32)
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = oracle.bf16_round(x)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_forward_exact_matches_torch():
    p = rand_params(2)
    X, _ = rand_rows(3, 64)
    y = oracle.mlp_forward(p, X, exact=True)
    ref = torch_forward(p, X).numpy()
    assert np.max(np.abs(y - ref)) <= 1e-5 * (1 + np.max(np.abs(ref)))


def test_gradient_exact_matches_autograd():
    p = rand_params(4)
    X, T = rand_rows(5, 40)
    loss, g = oracle.mlp_grad(p, X, T, exact=True)
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    W1, b1, W2, b2, W3, b3 = (None,) * 6
    o = 0
    parts = []
    for shape in ((108, 64), (64,), (64, 64), (64,), (64, 3), (3,)):
        n = int(np.prod(shape))
        parts.append(pt[o:o + n].reshape(shape))
        o += n
    W1, b1, W2, b2, W3, b3 = parts
    x = torch.tensor(X[:, :108], dtype=torch.float64)
    gl = torch.nn.functional.gelu
    y = gl(gl(x @ W1 + b1, approximate="tanh") @ W2 + b2, approximate="tanh") @ W3 + b3
    L = ((y - torch.tensor(T, dtype=torch.float64)) ** 2).sum() / X.shape[0]
    L.backward()
    assert abs(loss - L.item()) <= 1e-6 * abs(L.item())
    ref = pt.grad.numpy()
    assert np.linalg.norm(g - ref) <= 1e-6 * np.linalg.norm(ref)


def test_gradient_exact_finite_differences():
    p = rand_params(6).astype(np.float64)
    X, T = rand_rows(7, 12)
    _, g = oracle.mlp_grad(p.astype(np.float32), X, T, exact=True)
    rng = np.random.default_rng(8)
    # central differences in float32 parameters: pick large-gradient entries
    idx = np.argsort(-np.abs(g))[:10].tolist() + rng.integers(0, NP, 10).tolist()
    p32 = p.astype(np.float32)
    for i in idx:
        h = np.float32(1e-2) * max(np.float32(1.0), abs(p32[i]))
        a, b = p32.copy(), p32.copy()
        a[i] += h
        b[i] -= h
        la, _ = oracle.mlp_grad(a, X, T, exact=True)
        lb, _ = oracle.mlp_grad(b, X, T, exact=True)
        fd = (la - lb) / (float(a[i]) - float(b[i]))
        assert abs(fd - g[i]) <= 2e-3 * (abs(g[i]) + 1e-3), (i, fd, g[i])


def test_bf16_mode_is_a_perturbation_of_exact():
    p = rand_params(9)
    X, T = rand_rows(10, 256)
    ye = oracle.mlp_forward(p, X, exact=True)
    yb = oracle.mlp_forward(p, X, exact=False)
    assert np.linalg.norm(yb - ye) <= 0.05 * np.linalg.norm(ye)
    _, ge = oracle.mlp_grad(p, X, T, exact=True)
    _, gb = oracle.mlp_grad(p, X, T, exact=False)
    assert np.linalg.norm(gb - ge) <= 0.05 * np.linalg.norm(ge)


def test_adam_matches_torch():
    rng = np.random.default_rng(11)
    p = rng.standard_normal(NP).astype(np.float32)
    m = np.zeros(NP, np.float32)
    v = np.zeros(NP, np.float32)
    pt = torch.tensor(p.copy(), requires_grad=True)
    opt = torch.optim.Adam([pt], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    for t in range(1, 6):
        g = rng.standard_normal(NP).astype(np.float32)
        oracle.mlp_adam(p, m, v, g, 1e-3, t)
        opt.zero_grad()
        pt.grad = torch.tensor(g)
        opt.step()
    assert np.max(np.abs(p - pt.detach().numpy())) <= 1e-6


def test_features_brute_force():
    n = (4, 3, 5)
    nx, ny, nz = n[0] + 1, n[1] + 1, n[2] + 1
    nn = nx * ny * nz
    rng = np.random.default_rng(12)
    mass = rng.uniform(0.5, 2.0, nn)
    vel = rng.standard_normal((3, nn))
    nodes = np.array([0, 7, nn // 2, nn - 1, nx * ny + nx + 1], dtype=np.int64)
    X = oracle.mlp_features(mass, vel, n, nodes, 2.0, 0.5)
    for r, q in enumerate(nodes):
        i, j, k = q % nx, (q // nx) % ny, q // (nx * ny)
        nb = 0
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    a, b, c = i + dx, j + dy, k + dz
                    want = np.zeros(4)
                    if 0 <= a < nx and 0 <= b < ny and 0 <= c < nz:
                        p = a + nx * (b + ny * c)
                        want = np.array([mass[p] / 2.0] +
                                        [mass[p] * vel[d, p] / (2.0 * 0.5) for d in range(3)])
                    assert np.allclose(X[r, 4 * nb:4 * nb + 4], want, rtol=1e-6, atol=0), (q, nb)
                    nb += 1
        assert not X[r, 108:].any()
