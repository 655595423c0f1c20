"""Oracle (test infrastructure only): numpy restatement of the learned
warm-start network of the HIP library (cuda_sto_pf_amd/csrc/pf_nn.hpp):
forward pass, mean-square loss and its gradients.

It is the checker for the GPU forward and backward kernels, never the thing
measured or shipped; only ``tests/`` may import it.  The network has no
counterpart in the reference.

Network: 5 periodic 3x3x3 convolutions with 16 channels (input channels 0..6
used, output channels 0..2 used), leaky ReLU (slope 0.1) after the first
four, linear output.  Parameters: per layer W[27][16][16] (tap = 9(dx+1) +
3(dy+1) + (dz+1), [cin][cout]) then bias[16].

With ``bf16=True`` the oracle rounds where the GPU stores BF16 (input
features, weights, hidden activations, the backward signals) and keeps every
sum in float64, so GPU and oracle differ only by FP32 accumulation order.
With ``bf16=False`` it is plain float64, which the finite-difference pin of
tests/test_oracle_nn.py checks.
"""
import numpy as np

CH, TAPS, LAYERS, IN_CH, OUT_CH = 16, 27, 5, 7, 3
W_PER_LAYER = TAPS * CH * CH
P_PER_LAYER = W_PER_LAYER + CH
P_TOTAL = LAYERS * P_PER_LAYER
SLOPE = 0.1


def bf16(x):
    """Round to BF16 (nearest even) and return as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


def offsets():
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(TAPS)]


def layer_params(params, L, round_w):
    base = L * P_PER_LAYER
    W = np.asarray(params[base:base + W_PER_LAYER], dtype=np.float64).reshape(TAPS, CH, CH)
    b = np.asarray(params[base + W_PER_LAYER:base + P_PER_LAYER], dtype=np.float64)
    return (bf16(W) if round_w else W), b


def conv(X, W):
    """Periodic correlation: out[v][n] = sum_t sum_k X[v + off(t)][k] W[t][k][n];
    X is [B, Nx, Ny, Nz, 16]."""
    out = np.zeros(X.shape[:4] + (W.shape[2],))
    for t, (dx, dy, dz) in enumerate(offsets()):
        Xs = np.roll(X, shift=(-dx, -dy, -dz), axis=(1, 2, 3))
        out += Xs @ W[t]
    return out


def conv_transpose(G, W):
    """Adjoint of conv: out[v][k] = sum_t sum_n G[v - off(t)][n] W[t][k][n]."""
    out = np.zeros(G.shape[:4] + (W.shape[1],))
    for t, (dx, dy, dz) in enumerate(offsets()):
        Gs = np.roll(G, shift=(dx, dy, dz), axis=(1, 2, 3))
        out += Gs @ W[t].T
    return out


def forward(params, in7, shape, bf=True):
    """in7: [B][S][7] -> (output [B][3][S], activations)."""
    B = in7.shape[0]
    X = np.zeros((B,) + tuple(shape) + (CH,))
    X[..., :IN_CH] = np.asarray(in7, dtype=np.float64).reshape((B,) + tuple(shape) + (IN_CH,))
    if bf:
        X = bf16(X)
    acts = [X]
    for L in range(LAYERS):
        W, b = layer_params(params, L, bf)
        z = conv(acts[-1], W) + b
        if L < LAYERS - 1:
            a = np.where(z > 0, z, SLOPE * z)
            acts.append(bf16(a) if bf else a)
        else:
            y = z[..., :OUT_CH]
    S = int(np.prod(shape))
    return np.moveaxis(y.reshape(B, S, OUT_CH), 2, 1), acts


def loss_and_gradients(params, in7, target, shape, bf=True):
    """Mean-square loss over [B][3][S] and its gradient with respect to all
    P_TOTAL parameters."""
    B = in7.shape[0]
    S = int(np.prod(shape))
    y, acts = forward(params, in7, shape, bf)
    d = y - np.asarray(target, dtype=np.float64)
    count = d.size
    loss = float(np.sum(d * d) / count)
    G = np.zeros((B,) + tuple(shape) + (CH,))
    g3 = 2.0 * d / count                      # [B][3][S]
    G[..., :OUT_CH] = np.moveaxis(g3, 1, 2).reshape((B,) + tuple(shape) + (OUT_CH,))
    if bf:
        G = bf16(G)
    grad = np.zeros(P_TOTAL)
    for L in range(LAYERS - 1, -1, -1):
        X = acts[L]
        base = L * P_PER_LAYER
        dW = np.zeros((TAPS, CH, CH))
        for t, (dx, dy, dz) in enumerate(offsets()):
            Xs = np.roll(X, shift=(-dx, -dy, -dz), axis=(1, 2, 3))
            dW[t] = np.einsum("bxyzk,bxyzn->kn", Xs, G)
        grad[base:base + W_PER_LAYER] = dW.ravel()
        grad[base + W_PER_LAYER:base + P_PER_LAYER] = G.sum(axis=(0, 1, 2, 3))
        if L > 0:
            W, _ = layer_params(params, L, bf)
            dX = conv_transpose(G, W)
            dX = np.where(X > 0, dX, SLOPE * dX)
            G = bf16(dX) if bf else dX
    return loss, grad


def init_params(seed, in_scale=1.0):
    """He-style initialisation of the used channels; unused input channels of
    layer 0 and output channels >= 3 of the last layer are zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(P_TOTAL, dtype=np.float32)
    for L in range(LAYERS):
        cin = IN_CH if L == 0 else CH
        cout = OUT_CH if L == LAYERS - 1 else CH
        std = np.sqrt(2.0 / (TAPS * cin)) * (in_scale if L == 0 else 1.0)
        if L == LAYERS - 1:
            std *= 0.1
        W = np.zeros((TAPS, CH, CH), dtype=np.float32)
        W[:, :cin, :cout] = rng.normal(0.0, std, (TAPS, cin, cout))
        base = L * P_PER_LAYER
        p[base:base + W_PER_LAYER] = W.ravel()
    return p
