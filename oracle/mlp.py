"""CPU oracle of the learned grid-velocity corrector (test infrastructure only:
tests and bench's CPU baseline compare the product against it; the product
never calls it).

The corrector has no counterpart in the reference; this module restates the
network, its loss and gradient, and the per-node feature layout from their
definitions in water-sim-cuda_amd/csrc/corrector.hip, in float64 (exact
arithmetic up to rounding) and with optional BF16 rounding at the points where
the GPU kernels round (inputs, weights, activations, back-propagated
gradients), so the GPU results can be compared at their own precision.
The analytic gradient is pinned by central finite differences in float64
(tests/test_corrector_oracle.py).
"""
from __future__ import annotations

import numpy as np

NF, STRIDE, NH, NO = 66, 72, 64, 3
SHAPES = [("W1", (NF, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NH)), ("b3", (NH,)),
          ("W4", (NH, NO)), ("b4", (NO,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SHAPES])


def init_params(seed: int = 0, scale: float = 1.0) -> np.ndarray:
    """He-style initialisation (float32)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            d[name] = rng.normal(0.0, scale * np.sqrt(2.0 / shape[0]), size=shape)
        else:
            d[name] = np.zeros(shape)
    return pack(d).astype(np.float32)


def bf16(a) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even), as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    if nan.any():
        out = out.copy()
        out[nan] = np.nan
    return out


def _q(a, on: bool):
    return bf16(a).astype(np.float64) if on else np.asarray(a, dtype=np.float64)


def forward(X, params, mean, inv_std, out_scale=1.0, emulate_bf16=True, cache=False, dead=0.0):
    """Corrections (n, 3) of feature rows X (n, 66 or 72); rows whose output
    magnitude (before out_scale) is below `dead` give zero."""
    X = np.asarray(X, dtype=np.float32)[:, :NF]
    W = unpack(np.asarray(params, dtype=np.float32))
    xn = ((X - np.asarray(mean, np.float32)[None, :]) * np.asarray(inv_std, np.float32)[None, :]).astype(np.float32)
    q = emulate_bf16
    h0 = _q(xn, q)
    a1 = h0 @ _q(W["W1"], q) + W["b1"].astype(np.float64)
    h1 = _q(np.maximum(a1, 0.0).astype(np.float32), q) if q else np.maximum(a1, 0.0)
    a2 = h1 @ _q(W["W2"], q) + W["b2"].astype(np.float64)
    h2 = _q(np.maximum(a2, 0.0).astype(np.float32), q) if q else np.maximum(a2, 0.0)
    a3 = h2 @ _q(W["W3"], q) + W["b3"].astype(np.float64)
    h3 = _q(np.maximum(a3, 0.0).astype(np.float32), q) if q else np.maximum(a3, 0.0)
    y = h3 @ _q(W["W4"], q) + W["b4"].astype(np.float64)
    if cache:
        return y, (h0, h1, h2, h3)
    if dead > 0.0:
        y = np.where((np.sqrt((y * y).sum(1)) < dead)[:, None], 0.0, y)
    return y * out_scale


def loss_and_grad(X, T, params, mean, inv_std, out_scale=1.0, emulate_bf16=False):
    """L = mean over rows and outputs of (y - T / out_scale)^2 and dL/dparams
    (float64), back-propagated analytically."""
    q = emulate_bf16
    W = unpack(np.asarray(params, dtype=np.float32))
    y, (h0, h1, h2, h3) = forward(X, params, mean, inv_std, 1.0, emulate_bf16=q, cache=True)
    t = np.asarray(T, dtype=np.float64) / out_scale
    n = X.shape[0]
    d = y - t
    L = float((d * d).sum() / (3 * n))
    g4 = 2.0 * d / (3 * n)
    gq = (lambda a: _q(a.astype(np.float32), True)) if q else (lambda a: a)
    grads = {"W4": h3.T @ gq(g4), "b4": g4.sum(0)}
    g3 = (gq(g4) @ _q(W["W4"], q).T) * (h3 > 0)
    grads["W3"] = h2.T @ gq(g3)
    grads["b3"] = g3.sum(0)
    g2 = (gq(g3) @ _q(W["W3"], q).T) * (h2 > 0)
    grads["W2"] = h1.T @ gq(g2)
    grads["b2"] = g2.sum(0)
    g1 = (gq(g2) @ _q(W["W2"], q).T) * (h1 > 0)
    grads["W1"] = h0.T @ gq(g1)
    grads["b1"] = g1.sum(0)
    return L, pack(grads)


def loss(X, T, params, mean, inv_std, out_scale=1.0):
    y = forward(X, params, mean, inv_std, 1.0, emulate_bf16=False)
    d = y - np.asarray(T, dtype=np.float64) / out_scale
    return float((d * d).sum() / (3 * X.shape[0]))


# ---- per-node features (csrc/corrector.hip, node_features) ----

def node_features(grid4, sdf, cfg, x: int, y: int, z: int, dt: float) -> np.ndarray:
    """Feature row (72 floats) of node (x, y, z) from the grid after the grid
    update (planes mass, vx, vy, vz; index z*X*Y + y*X + x) and the node SDF."""
    sx, sy, sz = cfg.SIZE_X, cfg.SIZE_Y, cfg.SIZE_Z
    m, vx, vy, vz = (np.asarray(p, dtype=np.float32).reshape(sz, sy, sx) for p in grid4)
    s = np.asarray(sdf, dtype=np.float32).reshape(sz, sy, sx)
    f = np.zeros(STRIDE, dtype=np.float64)
    f[0] = m[z, y, x] / 8.0
    f[1:4] = np.array([vx[z, y, x], vy[z, y, x], vz[z, y, x]], dtype=np.float64) * dt
    f[4] = min(max(float(s[z, y, x]), -4.0), 4.0) / 4.0
    # getNormalAtNode, common.cuh:195-203
    n = np.array([s[z, y, min(x + 1, sx - 1)] - s[z, y, max(x - 1, 0)],
                  s[z, min(y + 1, sy - 1), x] - s[z, max(y - 1, 0), x],
                  s[min(z + 1, sz - 1), y, x] - s[max(z - 1, 0), y, x]], dtype=np.float64)
    f[5:8] = n / (np.sqrt((n * n).sum()) + 1e-9)
    o = 0
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                xx, yy, zz = x + dx, y + dy, z + dz
                if 0 <= xx < sx and 0 <= yy < sy and 0 <= zz < sz:
                    mm = float(m[zz, yy, xx])
                    sp = np.sqrt(float(vx[zz, yy, xx]) ** 2 + float(vy[zz, yy, xx]) ** 2 + float(vz[zz, yy, xx]) ** 2)
                    f[8 + o] = mm / 8.0
                    f[35 + o] = mm * sp * dt / 8.0
                o += 1
    f[62] = cfg.GRAVITY * dt * dt
    f[63] = cfg.COMPRESSION * dt * dt
    p = float(cfg.PADDING)
    w = min(x - p, sx - 1 - p - x, y - p, sy - 1 - p - y, z - p, sz - 1 - p - z)
    f[64] = min(max(w / 8.0, 0.0), 1.0)
    f[65] = y / sy
    return f
