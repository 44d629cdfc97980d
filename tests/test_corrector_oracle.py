"""CPU pins of the corrector's oracle (oracle/mlp.py): the analytic gradient
against central finite differences, BF16 rounding against hand-worked
values, the forward pass against closed forms, node features against a
hand-built grid."""
import numpy as np
import pytest

from oracle import mlp


def test_bf16_rounding_known_values():
    one = np.float32(1.0)
    e7 = np.float32(2.0 ** -7)
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, -2.5, 0.0, 3.0e38],
                     dtype=np.float32)
    got = mlp.bf16(vals)
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6; above the tie -> up
    np.testing.assert_array_equal(got[:4], np.array([one, one, one + 2 * e7, one + e7], dtype=np.float32))
    assert got[4] == -2.5 and got[5] == 0.0
    assert np.isnan(mlp.bf16(np.array([np.nan], np.float32)))[0]


def test_forward_closed_forms():
    d = mlp.unpack(np.zeros(mlp.NPARAM))
    d["b4"] = np.array([1.0, -2.0, 3.0])
    p = mlp.pack(d).astype(np.float32)
    X = np.random.default_rng(0).normal(size=(5, 66)).astype(np.float32)
    y = mlp.forward(X, p, np.zeros(66), np.ones(66), out_scale=2.0)
    np.testing.assert_allclose(y, np.tile([2.0, -4.0, 6.0], (5, 1)))
    # a single ReLU path: y_1 = 2 relu(relu(relu(x_0)))
    d = mlp.unpack(np.zeros(mlp.NPARAM))
    d["W1"][0, 0] = d["W2"][0, 0] = d["W3"][0, 0] = 1.0
    d["W4"][0, 1] = 2.0
    p = mlp.pack(d).astype(np.float32)
    X = np.zeros((3, 66), np.float32)
    X[:, 0] = [-1.0, 0.5, 2.0]
    y = mlp.forward(X, p, np.zeros(66), np.ones(66), emulate_bf16=False)
    np.testing.assert_allclose(y[:, 1], [0.0, 1.0, 4.0])
    np.testing.assert_allclose(y[:, [0, 2]], 0.0)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    n = 37
    X = rng.normal(size=(n, 66)).astype(np.float32)
    T = rng.normal(size=(n, 3)).astype(np.float32)
    p = mlp.init_params(5).astype(np.float64)
    p[mlp.NPARAM - 3:] = 0.1
    mean, inv = np.zeros(66), np.ones(66)
    L, g = mlp.loss_and_grad(X, T, p, mean, inv, emulate_bf16=False)
    assert abs(L - mlp.loss(X, T, p, mean, inv)) < 1e-12
    idx = np.concatenate([rng.choice(mlp.NPARAM, 80, replace=False), np.arange(mlp.NPARAM - 6, mlp.NPARAM)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        # float64 forward: evaluate with float64 params directly
        fd = (_loss64(X, T, pp) - _loss64(X, T, pm)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (i, fd, g[i])


def _loss64(X, T, p):
    d = mlp.unpack(p)
    h = X.astype(np.float64)
    for k in (1, 2, 3):
        h = np.maximum(h @ d[f"W{k}"] + d[f"b{k}"], 0.0)
    y = h @ d["W4"] + d["b4"]
    r = y - T
    return float((r * r).sum() / (3 * X.shape[0]))


def test_gradient_bf16_emulation_close_to_exact():
    rng = np.random.default_rng(4)
    X = rng.normal(size=(256, 66)).astype(np.float32)
    T = rng.normal(size=(256, 3)).astype(np.float32)
    p = mlp.init_params(6)
    _, g = mlp.loss_and_grad(X, T, p, np.zeros(66), np.ones(66), emulate_bf16=False)
    _, gq = mlp.loss_and_grad(X, T, p, np.zeros(66), np.ones(66), emulate_bf16=True)
    assert np.linalg.norm(gq - g) / np.linalg.norm(g) < 3e-2


class _Cfg:
    SIZE_X = SIZE_Y = SIZE_Z = 8
    PADDING = 2
    GRAVITY = 9.8
    COMPRESSION = 10.0


def test_node_features_hand_grid():
    cfg = _Cfg()
    G = 8 ** 3
    m = np.zeros(G, np.float32)
    v = np.zeros((3, G), np.float32)
    idx = lambda x, y, z: z * 64 + y * 8 + x
    m[idx(4, 4, 4)] = 8.0
    v[:, idx(4, 4, 4)] = [1.0, -2.0, 2.0]
    m[idx(5, 4, 4)] = 4.0
    v[:, idx(5, 4, 4)] = [0.0, 3.0, 4.0]
    sdf = np.tile(np.arange(8, dtype=np.float32)[None, :, None] - 3.0, (8, 1, 8)).reshape(-1)  # y - 3
    f = mlp.node_features([m, *v], sdf, cfg, 4, 4, 4, dt=0.1)
    assert f[0] == 1.0
    np.testing.assert_allclose(f[1:4], [0.1, -0.2, 0.2], rtol=1e-6)
    assert abs(f[4] - 0.25) < 1e-7
    np.testing.assert_allclose(f[5:8], [0.0, 1.0, 0.0], atol=1e-6)
    centre = 13  # (0, 0, 0) offset, x fastest
    assert f[8 + centre] == 1.0 and f[8 + centre + 1] == 0.5
    np.testing.assert_allclose(f[35 + centre], 8.0 * 3.0 * 0.1 / 8.0, rtol=1e-6)
    np.testing.assert_allclose(f[35 + centre + 1], 4.0 * 5.0 * 0.1 / 8.0, rtol=1e-6)
    np.testing.assert_allclose(f[62:64], [9.8 * 0.01, 10.0 * 0.01], rtol=1e-6)
    assert f[64] == pytest.approx(min(2, 1, 2, 1, 2, 1) / 8.0)
    assert f[65] == 0.5
    assert not f[66:].any()


def test_dead_zone_drops_small_outputs():
    """forward(dead=d) zeroes exactly the rows whose output norm (before
    out_scale) is below d and leaves the others as they are."""
    rng = np.random.default_rng(3)
    X = rng.normal(size=(500, 72)).astype(np.float32)
    p = mlp.init_params(4)
    y = mlp.forward(X, p, np.zeros(66), np.ones(66), 1.0, emulate_bf16=False)
    norm = np.sqrt((y * y).sum(1))
    d = float(np.median(norm))
    z = mlp.forward(X, p, np.zeros(66), np.ones(66), 3.0, emulate_bf16=False, dead=d)
    small = norm < d
    assert small.any() and (~small).any()
    assert not z[small].any()
    np.testing.assert_allclose(z[~small], 3.0 * y[~small], rtol=1e-12)
