"""Pins of the neural-field oracle (oracle/neural.py), which checks the GPU
training and inference kernels: trilinear hash-grid encoding reproduces
linear fields on densely indexed levels, the spatial hash matches integer
arithmetic, and the backward pass matches finite differences of the forward
pass (binary64, so that a dropped term, a transposed product or a wrong mask
shows up as a gradient error)."""
import numpy as np
import pytest

from oracle import neural as NR


def small_model(seed):
    rng = np.random.default_rng(seed)
    res = NR.level_resolutions(16, 256)
    table = (rng.standard_normal((NR.L_LEVELS, NR.T_SIZE, 2)) * 0.5).astype(np.float16)
    W1 = (rng.standard_normal((32, 64)) * 0.4).astype(np.float16)
    W2 = (rng.standard_normal((64, 64)) * 0.25).astype(np.float16)
    W3 = (rng.standard_normal(64) * 0.2).astype(np.float32)
    return res, table, W1, W2, W3


def test_level_resolutions_geometric():
    r = NR.level_resolutions(16, 2048)
    assert r[0] == 16 and r[-1] == 2048 and np.all(np.diff(r.astype(np.int64)) > 0)
    b = np.exp((np.log(2048) - np.log(16)) / 15)
    np.testing.assert_array_equal(r, np.floor(16 * b ** np.arange(16)).astype(np.uint32))


def test_hash_matches_integer_arithmetic():
    res = 1000  # (res+1)^3 > 2^19: hashed level
    cx, cy, cz = (np.array(v, np.uint32) for v in ([0, 1, 999, 12345], [0, 7, 3, 54321], [0, 2, 1000, 77]))
    got = NR.corner_index(res, cx, cy, cz)
    for i in range(4):
        h = (int(cx[i]) ^ ((int(cy[i]) * 2654435761) % 2 ** 32) ^ ((int(cz[i]) * 805459861) % 2 ** 32)) % 2 ** 19
        assert got[i] == h
    # dense level: row-major with side res + 1
    assert NR.corner_index(16, np.array([3], np.uint32), np.array([5], np.uint32), np.array([7], np.uint32))[0] == \
        3 + 17 * (5 + 17 * 7)


def test_encoding_reproduces_linear_field_on_dense_levels():
    res = NR.level_resolutions(16, 2048)
    table = np.zeros((NR.L_LEVELS, NR.T_SIZE, 2), np.float16)
    l = 0
    side = int(res[l]) + 1
    k = np.arange(side ** 3)
    cx, cy, cz = k % side, (k // side) % side, k // (side * side)
    table[l, :side ** 3, 0] = (cx + 2 * cy + 3 * cz).astype(np.float16)
    table[l, :side ** 3, 1] = 1.0
    x = np.random.default_rng(3).random((500, 3), dtype=np.float32)
    for enc in (NR.encode(x, table, res)[0], NR.encode64(x, table, res)[0]):
        want = (x[:, 0] + 2 * x[:, 1] + 3 * x[:, 2]).astype(np.float64) * float(res[l])
        np.testing.assert_allclose(enc[:, 0], want, rtol=0, atol=2e-4)
        np.testing.assert_allclose(enc[:, 1], 1.0, atol=1e-6)
        assert np.all(enc[:, 2:] == 0)


def test_backward_matches_finite_differences():
    res, table, W1, W2, W3 = small_model(11)
    rng = np.random.default_rng(12)
    x = rng.random((48, 3))
    t = rng.standard_normal(48)
    y, g = NR.backward(x, t, table, W1, W2, W3, res, half=False)

    def loss(tb=table, w1=W1, w2=W2, w3=W3):
        yy, _ = NR.forward(x, tb, w1, w2, w3, res, half=False)
        return float(np.sum((yy - t) ** 2))

    eps = 1e-6
    checks = 0
    for name, arr, shape in (("W1", W1, (32, 64)), ("W2", W2, (64, 64)), ("W3", W3, (64,))):
        base = arr.astype(np.float64)
        for _ in range(8):
            idx = tuple(int(rng.integers(0, s)) for s in shape)
            if abs(g[name][idx]) < 1e-8:
                continue
            p, m = base.copy(), base.copy()
            p[idx] += eps
            m[idx] -= eps
            kw = {"w1": W1, "w2": W2, "w3": W3}
            kw[{"W1": "w1", "W2": "w2", "W3": "w3"}[name]] = p
            lp = loss(**kw)
            kw[{"W1": "w1", "W2": "w2", "W3": "w3"}[name]] = m
            lm = loss(**kw)
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[name][idx]) <= 1e-5 * max(1.0, abs(fd)), (name, idx, fd, g[name][idx])
            checks += 1
    # table entries actually touched by the points
    _, st = NR.forward(x, table, W1, W2, W3, res, half=False)
    tb64 = table.astype(np.float64)
    for l in (0, 5, 15):
        idx_l = st["cache"][l][0]
        for j in range(4):
            e = int(idx_l[j, j % 8])
            f = j % 2
            p, m = tb64.copy(), tb64.copy()
            p[l, e, f] += eps
            m[l, e, f] -= eps
            fd = (loss(tb=p) - loss(tb=m)) / (2 * eps)
            assert abs(fd - g["table"][l, e, f]) <= 1e-5 * max(1.0, abs(fd)), (l, e, f)
            checks += 1
    assert checks >= 30


def test_planted_backward_mistakes_are_caught():
    """The finite-difference bar fails a transposed W2 product and a dropped
    ReLU mask (mistakes planted in a copy of the oracle's backward)."""
    res, table, W1, W2, W3 = small_model(13)
    rng = np.random.default_rng(14)
    x = rng.random((32, 3))
    t = rng.standard_normal(32)
    _, g = NR.backward(x, t, table, W1, W2, W3, res, half=False)
    y, st = NR.forward(x, table, W1, W2, W3, res, half=False)
    dy = 2 * (y - t)
    G2 = (dy[:, None] * W3.astype(np.float64)[None, :]) * (st["H2"] > 0)
    wrong_T = (G2 @ W2.astype(np.float64)) * (st["H1"] > 0)           # W2 instead of W2^T
    no_mask = G2 @ W2.astype(np.float64).T                             # ReLU mask dropped
    right = (G2 @ W2.astype(np.float64).T) * (st["H1"] > 0)
    dW1_right = st["X"].T @ right
    np.testing.assert_allclose(dW1_right, g["W1"], rtol=1e-10, atol=1e-12)
    for bad in (wrong_T, no_mask):
        dW1_bad = st["X"].T @ bad
        assert np.linalg.norm(dW1_bad - g["W1"]) > 1e-2 * np.linalg.norm(g["W1"])
