"""Pins of the network oracle (CPU only): the float64 gradients against
central finite differences of the loss; BF16 rounding against its exact
definition; the periodic convolution against a direct sum."""
import numpy as np

from oracle import nn_ref as nr


def _data(shape=(4, 4, 6), B=2, seed=0):
    rng = np.random.default_rng(seed)
    S = int(np.prod(shape))
    x = rng.normal(0.0, 1.0, (B, S, nr.IN_CH)).astype(np.float32)
    t = rng.normal(0.0, 0.3, (B, nr.OUT_CH, S))
    return x, t


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 1e-3], dtype=np.float32)
    got = nr.bf16(vals)
    # ties to even: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2 ** -7 and got[3] == 1.0 + 2 ** -6
    assert got[4] == -2.5
    assert abs(got[5] - 1e-3) <= 2 ** -9 * 1e-3


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(1)
    X = rng.normal(size=(1, 3, 4, 5, nr.CH))
    W = rng.normal(size=(nr.TAPS, nr.CH, nr.CH))
    out = nr.conv(X, W)
    x, y, z = 2, 0, 4
    s = np.zeros(nr.CH)
    for t, (dx, dy, dz) in enumerate(nr.offsets()):
        s += X[0, (x + dx) % 3, (y + dy) % 4, (z + dz) % 5] @ W[t]
    assert np.allclose(out[0, x, y, z], s, rtol=1e-13, atol=1e-13)


def test_gradients_match_finite_differences():
    shape = (4, 4, 6)
    x, t = _data(shape)
    p = nr.init_params(3).astype(np.float64)
    rng = np.random.default_rng(5)
    p += rng.normal(0.0, 0.02, p.shape) * (p != 0)          # biases and weights away from init
    for L in range(nr.LAYERS):
        base = L * nr.P_PER_LAYER + nr.W_PER_LAYER
        p[base:base + (nr.OUT_CH if L == nr.LAYERS - 1 else nr.CH)] = rng.normal(0.0, 0.05)
    loss, g = nr.loss_and_gradients(p, x, t, shape, bf=False)
    idx = []
    for L in range(nr.LAYERS):
        nz = np.flatnonzero(p[L * nr.P_PER_LAYER:(L + 1) * nr.P_PER_LAYER]) + L * nr.P_PER_LAYER
        idx += list(rng.choice(nz, 6, replace=False))
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp, _ = nr.loss_and_gradients(q, x, t, shape, bf=False)
        q[i] -= 2 * h
        lm, _ = nr.loss_and_gradients(q, x, t, shape, bf=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(fd), 1e-3), (i, fd, g[i])
