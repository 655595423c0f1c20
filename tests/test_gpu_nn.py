"""Learned warm start on the GPU: the library's forward and backward kernels
(BF16 MFMA, FP32 accumulation) against the numpy oracle that rounds at the
same points (oracle/nn_ref.py); training through pf_nn_collect lowers the
loss; an attached network and a caller-supplied start leave converged
results unchanged."""
import ctypes as C

import numpy as np
import pytest

from cuda_sto_pf_amd.configs import config
from oracle import nn_ref as nr


def _params(seed):
    p = nr.init_params(seed).astype(np.float64)
    rng = np.random.default_rng(seed + 100)
    p += rng.normal(0.0, 0.02, p.shape) * (p != 0)
    for L in range(nr.LAYERS):
        base = L * nr.P_PER_LAYER + nr.W_PER_LAYER
        n = nr.OUT_CH if L == nr.LAYERS - 1 else nr.CH
        p[base:base + n] = rng.normal(0.0, 0.05, n)
    return p.astype(np.float32)


def _inputs(shape, B, seed):
    rng = np.random.default_rng(seed)
    S = int(np.prod(shape))
    x = rng.normal(0.0, 1.0, (B, S, nr.IN_CH)).astype(np.float32)
    t = rng.normal(0.0, 0.2, (B, nr.OUT_CH, S)).astype(np.float32)
    return x, t


SHAPES = [(8, 8, 32), (6, 5, 20)]


@pytest.mark.gpu
@pytest.mark.parametrize("shape", SHAPES, ids=["8x8x32", "6x5x20"])
def test_forward_matches_oracle(pf, shape):
    p = _params(1)
    net = pf.Net(shape, max_batch=2)
    net.set_params(p)
    x, _ = _inputs(shape, 2, 2)
    got = net.forward(x, shape)
    ref, _ = nr.forward(p, x, shape, bf=True)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"forward {shape}: rel={rel:.2e} max|d|={np.abs(got - ref).max():.2e} scale={np.abs(ref).max():.2e}",
          flush=True)
    assert rel < 5e-3
    assert np.abs(got - ref).max() < 3e-2 * np.abs(ref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("shape", SHAPES, ids=["8x8x32", "6x5x20"])
def test_gradients_match_oracle(pf, shape):
    p = _params(4)
    net = pf.Net(shape, max_batch=2)
    net.set_params(p)
    x, t = _inputs(shape, 2, 5)
    loss, g = net.gradients(x, t, shape)
    lref, gref = nr.loss_and_gradients(p, x, t, shape, bf=True)
    print(f"loss gpu={loss:.8e} oracle={lref:.8e}", flush=True)
    assert abs(loss - lref) <= 1e-4 * lref
    for L in range(nr.LAYERS):
        sl = slice(L * nr.P_PER_LAYER, (L + 1) * nr.P_PER_LAYER)
        rel = np.linalg.norm(g[sl] - gref[sl]) / np.linalg.norm(gref[sl])
        print(f"layer {L}: |g|={np.linalg.norm(gref[sl]):.3e} rel err={rel:.2e}", flush=True)
        assert rel < 1e-2


@pytest.mark.gpu
def test_training_on_solver_snapshots_lowers_loss(pf):
    n = 32
    pd, phi, xB, meta = config(2, n)
    sim = pf.Simulation(pd, pf.PF_MIXED)
    sim.set_fields(phi, xB, xB_cap=meta["xB_max_safe"])
    net = pf.Net((n, n, n), max_batch=4, capacity=6)
    net.set_scales(1.0, 0.1, 0.05)
    net.set_params(nr.init_params(7))
    for k in range(6):
        sim.step(2)
        assert net.collect(sim) == k + 1
    idx = np.arange(6)
    l0 = net.eval_loss(idx[:4])
    losses = [net.train_step(idx[(i % 2) * 2:(i % 2) * 2 + 4], 2e-3) for i in range(80)]
    l1 = net.eval_loss(idx[:4])
    print(f"training loss {l0:.4e} -> {l1:.4e} (last steps {losses[-3:]})", flush=True)
    assert np.isfinite(l1) and l1 < 0.5 * l0


@pytest.mark.gpu
@pytest.mark.parametrize("prec,pool", [("fp64", 1), ("mixed", 1), ("fp64", 2), ("mixed", 2)])
def test_attached_network_leaves_results_unchanged(pf, prec, pool):
    """A network at the cold start (on the grid itself, or on its 2x2x2
    averages with the prediction read back at every voxel) moves only where
    the fixed point starts: converged x_B agrees to the solver tolerance."""
    n = 32
    pd, phi, xB, meta = config(2, n)
    pd = dict(pd, elastic_warm_start_order=3, elastic_anderson_depth=3)
    P = pf.PF_FP64 if prec == "fp64" else pf.PF_MIXED
    a = pf.Simulation(pd, P)
    b = pf.Simulation(pd, P)
    for s in (a, b):
        s.set_fields(phi, xB, xB_cap=meta["xB_max_safe"])
    net = pf.Net((n // pool,) * 3, max_batch=1)
    net.set_scales(1.0, 0.1, 0.05)
    net.set_params(_params(9))
    net.set_pool(pool)
    b.attach_net(net)
    for _ in range(3):
        a.step()
        b.step()
    da = np.abs(a.field(pf.PF_XB) - b.field(pf.PF_XB)).max()
    print(f"{prec}: attached vs not, max|dxB|={da:.2e}", flush=True)
    assert da < 1e-7
    assert b.scalars()["last_elastic_residual"] <= 1e-8


@pytest.mark.gpu
def test_warm_start_callback_supplies_the_start(pf):
    """A caller start equal to the converged solution: the cold solve stops
    at the minimum update count."""
    n = 32
    pd, phi, xB, meta = config(2, n)
    ref = pf.Simulation(pd, pf.PF_FP64)
    ref.set_fields(phi, xB, xB_cap=meta["xB_max_safe"])
    ref.step()
    cold_iters = ref.scalars()["last_elastic_iterations"]
    u_conv = ref.elastic_state().copy()
    sim = pf.Simulation(pd, pf.PF_FP64)
    sim.set_fields(phi, xB, xB_cap=meta["xB_max_safe"])
    calls = []

    @C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p)
    def cb(h, user):
        calls.append(1)
        u = np.ascontiguousarray(u_conv)
        return pf.lib().pf_set_warm_start(h, u.ctypes.data_as(C.c_void_p))

    pf.lib().pf_set_warm_start_callback(sim._h, C.cast(cb, C.c_void_p), None)
    sim.step()
    it = sim.scalars()["last_elastic_iterations"]
    print(f"cold start {cold_iters} updates, supplied start {it}", flush=True)
    assert len(calls) == 1 and it <= 2 < cold_iters
    assert np.abs(sim.field(pf.PF_XB) - ref.field(pf.PF_XB)).max() < 1e-9


@pytest.mark.gpu
def test_pooled_collection_averages_the_full_resolution_snapshot(pf):
    """pf_nn_collect of a pooled network stores the 2x2x2 means of what a
    full-resolution network collects from the same state: the loss of the
    zero network (the mean square target) of the pooled snapshot lies between
    0 and that of the full one (averaging cannot raise a mean square), and the
    pooled network's predicted start reaches the same converged solution."""
    n = 32
    pd, phi, xB, meta = config(2, n)
    pd = dict(pd, elastic_warm_start_order=3, elastic_anderson_depth=3)
    losses = {}
    for pool in (1, 2):
        s = pf.Simulation(pd, pf.PF_MIXED)
        s.set_fields(phi, xB, xB_cap=meta["xB_max_safe"])
        net = pf.Net((n // pool,) * 3, max_batch=1, capacity=1)
        net.set_pool(pool)
        net.set_scales(1.0, 0.1, 1.0)
        net.collect(s)
        net.set_params(np.zeros(34640, dtype=np.float32))   # the zero network: loss = mean square target
        losses[pool] = net.eval_loss([0])
    print(f"mean square target: full {losses[1]:.4e}, pooled {losses[2]:.4e}", flush=True)
    assert 0.0 < losses[2] <= losses[1] * (1 + 1e-6)
    assert losses[2] > 0.05 * losses[1]   # the interface-scale force density survives the averaging
