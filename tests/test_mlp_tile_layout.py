"""The register algebra of the learned network's MFMA tile code
(exagoop_amd/csrc/mpm_mlp.hip), on the CPU.

A model of v_mfma_f32_32x32x16_bf16 in the lane layouts the kernel's header
states (A: lane l holds row l & 31, k = 8 (l >> 5) + j; B: lane l holds
column l & 31, same k; C: lane l holds column l & 31, register r row
8 (r >> 2) + 4 (l >> 5) + (r & 3)) runs the kernel's tile steps in FP64:
C tiles taken 8 registers at a time as K = rows operands, transposes by the
permuted identity, the swap23-permuted weight copies.  Forward output and
every gradient block must equal the plain matrix products, with each
gradient entry written exactly once; a swapped identity or an unpermuted
weight copy must not.  (The BF16 rounding is left out: the kernel's
transposes multiply by 1 and add zeros, so they are exact in any precision.)
"""
import numpy as np
import pytest

L = 64


def crow(r, lane):
    return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)


def swap23(s):
    return (s & ~12) | ((s << 1) & 8) | ((s >> 1) & 4)


def mfma(a, b, c):
    A = np.zeros((32, 16))
    B = np.zeros((16, 32))
    for lane in range(L):
        for j in range(8):
            A[lane & 31, 8 * (lane >> 5) + j] = a[lane][j]
            B[8 * (lane >> 5) + j, lane & 31] = b[lane][j]
    C = np.zeros((32, 32))
    for lane in range(L):
        for r in range(16):
            C[crow(r, lane), lane & 31] = c[lane][r]
    D = A @ B + C
    return np.array([[D[crow(r, lane), lane & 31] for r in range(16)] for lane in range(L)])


def zero16():
    return np.zeros((L, 16))


def half8(c, q):
    return c[:, 8 * q:8 * q + 8].copy()


def ident(q, perm):
    o = np.zeros((L, 8))
    for lane in range(L):
        n = lane & 31
        want = swap23(n) if perm else n
        base = 16 * q + 8 * (lane >> 5)
        for j in range(8):
            o[lane][j] = 1.0 if base + j == want else 0.0
    return o


def transpose(lo, hi, p0, p1):
    return mfma(hi, p1, mfma(lo, p0, zero16()))


def ld8(M, row_of, off_of):
    return np.array([M[row_of(lane), off_of(lane):off_of(lane) + 8] for lane in range(L)])


IN, INP, H, O = 108, 128, 64, 3


def operand_copies(W1, W2, W3, permute=True):
    """The BF16 operand copies of mpm_mlp.h (w1t, w2tp, w2p, w3tp, w3)."""
    sw = swap23 if permute else (lambda i: i)
    w1t = np.zeros((H, INP))
    w1t[:, :IN] = W1.T
    w2tp, w2p = np.zeros((H, H)), np.zeros((H, H))
    w3tp, w3 = np.zeros((32, H)), np.zeros((H, 16))
    for k in range(H):
        for n in range(H):
            w2tp[n, sw(k)] = W2[k, n]
            w2p[k, sw(n)] = W2[k, n]
        for n in range(O):
            w3tp[n, sw(k)] = W3[k, n]
            w3[k, n] = W3[k, n]
    return w1t, w2tp, w2p, w3tp, w3


def act(z):
    return np.tanh(z), 1.0 - np.tanh(z) ** 2  # any smooth activation: the algebra is the test


def tile(X, T, nrows, W, b, perm_ident=True, permute_weights=True):
    """The kernel's train_tile on one 32-row tile: Y (rows x 3) and the
    gradient entries {(block, i, j): value} with their write counts."""
    W1, W2, W3 = W
    b1, b2, b3 = b
    w1t, w2tp, w2p, w3tp, w3 = operand_copies(W1, W2, W3, permute_weights)
    col = np.arange(L) & 31
    p0, p1 = ident(0, perm_ident), ident(1, perm_ident)
    i0, i1 = ident(0, False), ident(1, False)
    xa = [ld8(X, lambda l: l & 31, lambda l, ks=ks: ks * 16 + 8 * (l >> 5)) for ks in range(8)]
    h1, h2, gd1, gd2, h1t, h2t = [None] * 2, [None] * 2, [None] * 2, [None] * 2, [None] * 2, [None] * 2
    for nt in range(2):
        acc = zero16()
        for ks in range(8):
            acc = mfma(xa[ks], ld8(w1t, lambda l: nt * 32 + (l & 31),
                                   lambda l: ks * 16 + 8 * (l >> 5)), acc)
        h, gd1[nt] = act(acc + b1[nt * 32 + col][:, None])
        h1[nt] = [half8(h, 0), half8(h, 1)]
        h1t[nt] = transpose(h1[nt][0], h1[nt][1], p0, p1)
    for nt in range(2):
        acc = zero16()
        for ks in range(4):
            acc = mfma(half8(h1t[ks >> 1], ks & 1),
                       ld8(w2tp, lambda l: nt * 32 + (l & 31), lambda l: ks * 16 + 8 * (l >> 5)),
                       acc)
        h, gd2[nt] = act(acc + b2[nt * 32 + col][:, None])
        h2[nt] = [half8(h, 0), half8(h, 1)]
        h2t[nt] = transpose(h2[nt][0], h2[nt][1], p0, p1)
    acc = zero16()
    for ks in range(4):
        acc = mfma(half8(h2t[ks >> 1], ks & 1),
                   ld8(w3tp, lambda l: l & 31, lambda l: ks * 16 + 8 * (l >> 5)), acc)
    y = acc + np.where(col < O, b3[np.minimum(col, O - 1)], 0.0)[:, None]
    Y = np.zeros((32, O))
    for lane in range(L):
        if (lane & 31) < O:
            for r in range(16):
                Y[crow(r, lane), lane & 31] = y[lane][r]
    grad, writes = {}, {}

    def put(key, v):
        grad[key] = v
        writes[key] = writes.get(key, 0) + 1

    dy = zero16()
    for lane in range(L):
        for r in range(16):
            row = crow(r, lane)
            if (lane & 31) < O and row < nrows:
                dy[lane][r] = 2.0 * (y[lane][r] - T[row, lane & 31]) / nrows
    dyb0, dyb1 = half8(dy, 0), half8(dy, 1)
    for m in range(2):
        g = mfma(h2[m][1], dyb1, mfma(h2[m][0], dyb0, zero16()))
        for lane in range(L):
            if (lane & 31) < O:
                for r in range(16):
                    put(("W3", m * 32 + crow(r, lane), lane & 31), g[lane][r])
    dya = half8(transpose(dyb0, dyb1, p0, p1), 0)
    dz2 = [None] * 2
    for nt in range(2):
        d = mfma(dya, ld8(w3, lambda l: nt * 32 + (l & 31), lambda l: 8 * (l >> 5)), zero16())
        d = d * gd2[nt]
        dz2[nt] = [half8(d, 0), half8(d, 1)]
        for lane in range(32):
            put(("b2", 0, nt * 32 + lane), d[lane].sum() + d[lane + 32].sum())
    for m in range(2):
        for nt in range(2):
            g = mfma(h1[m][1], dz2[nt][1], mfma(h1[m][0], dz2[nt][0], zero16()))
            for lane in range(L):
                for r in range(16):
                    put(("W2", m * 32 + crow(r, lane), nt * 32 + (lane & 31)), g[lane][r])
    dz2t = [transpose(dz2[nt][0], dz2[nt][1], p0, p1) for nt in range(2)]
    dz1 = [None] * 2
    for nt in range(2):
        acc = zero16()
        for ks in range(4):
            acc = mfma(half8(dz2t[ks >> 1], ks & 1),
                       ld8(w2p, lambda l: nt * 32 + (l & 31), lambda l: ks * 16 + 8 * (l >> 5)),
                       acc)
        d = acc * gd1[nt]
        dz1[nt] = [half8(d, 0), half8(d, 1)]
        for lane in range(32):
            put(("b1", 0, nt * 32 + lane), d[lane].sum() + d[lane + 32].sum())
    for m in range(4):
        xc = mfma(xa[2 * m + 1], i1, mfma(xa[2 * m], i0, zero16()))
        x0, x1 = half8(xc, 0), half8(xc, 1)
        for nt in range(2):
            g = mfma(x1, dz1[nt][1], mfma(x0, dz1[nt][0], zero16()))
            for lane in range(L):
                for r in range(16):
                    i = m * 32 + crow(r, lane)
                    if i < IN:
                        put(("W1", i, nt * 32 + (lane & 31)), g[lane][r])
    return Y, grad, writes


def reference(X, T, nrows, W, b):
    W1, W2, W3 = W
    b1, b2, b3 = b
    Z1 = X[:, :IN] @ W1 + b1
    H1, D1 = act(Z1)
    Z2 = H1 @ W2 + b2
    H2, D2 = act(Z2)
    Y = H2 @ W3 + b3
    dY = np.zeros_like(Y)
    dY[:nrows] = 2.0 * (Y[:nrows] - T[:nrows]) / nrows
    dZ2 = (dY @ W3.T) * D2
    dZ1 = (dZ2 @ W2.T) * D1
    return Y, {"W1": X[:, :IN].T @ dZ1, "W2": H1.T @ dZ2, "W3": H2.T @ dY,
               "b1": dZ1.sum(0)[None], "b2": dZ2.sum(0)[None]}


def dense(grad, name, shape):
    M = np.full(shape, np.nan)
    for (blk, i, j), v in grad.items():
        if blk == name:
            M[i, j] = v
    return M


def problem(seed, nrows=29):
    rng = np.random.default_rng(seed)
    W = (rng.standard_normal((IN, H)) * 0.2, rng.standard_normal((H, H)) * 0.2,
         rng.standard_normal((H, O)) * 0.2)
    b = (rng.standard_normal(H) * 0.1, rng.standard_normal(H) * 0.1, rng.standard_normal(O) * 0.1)
    X = np.zeros((32, INP))
    X[:nrows, :IN] = rng.standard_normal((nrows, IN))
    T = rng.standard_normal((32, O))
    return X, T, nrows, W, b


def test_tile_matches_matrix_products():
    X, T, nrows, W, b = problem(0)
    Y, grad, writes = tile(X, T, nrows, W, b)
    Yr, gr = reference(X, T, nrows, W, b)
    assert np.abs(Y - Yr).max() <= 1e-12 * np.abs(Yr).max()
    for name, ref in gr.items():
        M = dense(grad, name, ref.shape)
        assert not np.isnan(M).any(), name  # every entry written
        assert np.abs(M - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1.0), name
    assert set(writes.values()) == {1}  # and written once
    assert len(writes) == IN * H + H * H + H * O + 2 * H


@pytest.mark.parametrize("mistake", ["natural identity for the transposes",
                                     "unpermuted weight copies"])
def test_planted_layout_mistakes_fail(mistake):
    X, T, nrows, W, b = problem(1)
    kw = ({"perm_ident": False} if mistake.startswith("natural")
          else {"permute_weights": False})
    Y, grad, _ = tile(X, T, nrows, W, b, **kw)
    Yr, gr = reference(X, T, nrows, W, b)
    bad = np.abs(Y - Yr).max() > 1e-6 or any(
        np.abs(dense(grad, n, r.shape) - r).max() > 1e-6 for n, r in gr.items())
    assert bad
