"""Whole time step, HIP (libpfcore) against the FP64 CPU oracle on identical
seeded inputs.  FP64 mode must agree to 1e-10 in phi, Y and x_B; the mixed
mode to 1e-5 in x_B (the accuracy gate named for the mixed path)."""
import math

import numpy as np
import pytest

from cuda_sto_pf_amd.configs import base_params, composition, config, spheres_phi
from oracle.step_ref import Oracle, Params

FP64_TOL = 1e-10
MIXED_XB_TOL = 1e-5


def _oracle(pd, phi, xB, cap):
    return Oracle(Params(**pd), phi, xB, xB_cap=cap)


def _gpu(pf, pd, phi, xB, cap, prec):
    sim = pf.Simulation(pd, prec)
    sim.set_fields(phi, xB, xB_cap=cap)
    return sim


def _rect_case(shape=(16, 12, 8), hetero=True, zero_mode=True, elastic=True, seed=11):
    pd = base_params(shape, hetero=hetero, zero_mode=zero_mode, elastic=elastic)
    rng = np.random.default_rng(seed)
    phi = spheres_phi(shape, [(shape[0] / 2, shape[1] / 2, shape[2] / 2)], [3.0], width=1.5)
    xB = composition(phi, rng)
    return pd, phi, xB, 0.5


def _cases():
    out = []
    pd, phi, xB, meta = config(1, 16)
    out.append(("cfg1_16", pd, phi, xB, meta["xB_max_safe"]))
    pd, phi, xB, meta = config(2, 16)
    out.append(("cfg2_16", pd, phi, xB, meta["xB_max_safe"]))
    out.append(("rect_hetero_zm",) + _rect_case())
    out.append(("rect_hetero_nozm",) + _rect_case(zero_mode=False))
    out.append(("rect_noelastic",) + _rect_case(elastic=False))
    # Nz = 128 keeps 43 kz: spectral rows padded to whole 128-B lines
    out.append(("rect_kz_padded",) + _rect_case(shape=(24, 20, 128)))
    # length-256 axes: the plan of the 256^3 grid (two radix-16 stages, 16-column
    # y/x tiles, 86 retained kz in 96-wide padded rows) along z and along x
    out.append(("rect_z256",) + _rect_case(shape=(16, 16, 256)))
    out.append(("rect_x256",) + _rect_case(shape=(256, 12, 16)))
    # a fully anisotropic C and dC (every Voigt entry non-zero, both C and
    # C + dC positive definite): the general Voigt products, where the cubic
    # tensors of the workloads take the orthotropic ones
    pd, phi, xB, cap = _rect_case()
    a = np.random.default_rng(5).uniform(-1.0, 1.0, (6, 6))
    a = 0.5 * (a + a.T)
    pd = dict(pd, C=pd["C"] + 2.0 * a, dC=pd["dC"] - a)
    assert np.linalg.eigvalsh(pd["C"]).min() > 0 and np.linalg.eigvalsh(pd["C"] + pd["dC"]).min() > 0
    out.append(("rect_aniso", pd, phi, xB, cap))
    return out


CASES = _cases()


def _compare(o, sim, pf):
    phi = sim.field(pf.PF_PHI)
    Y = sim.field(pf.PF_Y)
    xB = sim.field(pf.PF_XB)
    return (np.abs(phi - o.phi).max(), np.abs(Y - o.Y).max(), np.abs(xB - o.xB).max())


@pytest.mark.gpu
@pytest.mark.parametrize("name,pd,phi,xB,cap", CASES, ids=[c[0] for c in CASES])
def test_fp64_step_parity(pf, name, pd, phi, xB, cap):
    nsteps = 5
    o = _oracle(pd, phi, xB, cap)
    sim = _gpu(pf, pd, phi, xB, cap, pf.PF_FP64)
    for s in range(nsteps):
        o.step()
        sim.step()
        sc = sim.scalars()
        dphi, dY, dx = _compare(o, sim, pf)
        print(f"{name} step {s + 1}: dphi={dphi:.2e} dY={dY:.2e} dxB={dx:.2e} "
              f"iters gpu={sc['last_elastic_iterations']} cpu={o.last_elastic_iterations} "
              f"res gpu={sc['last_elastic_residual']:.3e} cpu={o.last_elastic_residual:.3e}", flush=True)
        assert dphi < FP64_TOL and dY < FP64_TOL and dx < FP64_TOL
        if pd["elastic_enabled"]:
            assert sc["last_elastic_iterations"] == o.last_elastic_iterations
        if pd["zero_mode"]:
            # the device-side root find takes the oracle's steps to the same shift
            assert sc["zero_mode_iterations"] == o.zero_mode_iterations
            assert math.isclose(sc["zero_mode_lambda"], o.last_lambda, rel_tol=1e-9, abs_tol=1e-13)
    sc = sim.scalars()
    assert math.isclose(sc["mass_xBtot"], o.mass(), rel_tol=1e-12)
    if pd["zero_mode"]:
        assert abs(sc["mass_xBtot"] - o.target_mass) <= 1e-12 * max(abs(o.target_mass), 1.0)
    en = o.last_energies
    for key, gkey in (("grad", "f_grad"), ("dw", "f_dw"), ("elastic", "f_elastic"), ("chem", "f_chem")):
        assert math.isclose(sc[gkey], en[key], rel_tol=1e-10, abs_tol=1e-14), (key, sc[gkey], en[key])


@pytest.mark.gpu
@pytest.mark.parametrize("name,pd,phi,xB,cap", CASES[:3], ids=[c[0] for c in CASES[:3]])
def test_mixed_step_accuracy(pf, name, pd, phi, xB, cap):
    nsteps = 5
    o = _oracle(pd, phi, xB, cap)
    sim = _gpu(pf, pd, phi, xB, cap, pf.PF_MIXED)
    for s in range(nsteps):
        o.step()
        sim.step()
        dphi, dY, dx = _compare(o, sim, pf)
        print(f"mixed {name} step {s + 1}: dphi={dphi:.2e} dY={dY:.2e} dxB={dx:.2e}", flush=True)
    assert dx < MIXED_XB_TOL
    sc = sim.scalars()
    if pd["zero_mode"]:
        assert abs(sc["mass_xBtot"] - o.target_mass) <= 1e-12 * max(abs(o.target_mass), 1.0)


def _ext(pd, order=3, depth=3):
    q = dict(pd)
    q.update(elastic_warm_start_order=order, elastic_anderson_depth=depth)
    return q


EXT_CASES = [("cfg2_16_ext",) + (lambda c: (_ext(c[0]), c[1], c[2], c[3]["xB_max_safe"]))(config(2, 16)),
             ("rect_ext_m2",) + (lambda c: (_ext(c[0], 2, 2), c[1], c[2], c[3]))(_rect_case()),
             ("rect_kz_padded_ext",) + (lambda c: (_ext(c[0]), c[1], c[2], c[3]))(_rect_case(shape=(24, 20, 128)))]


@pytest.mark.gpu
@pytest.mark.parametrize("name,pd,phi,xB,cap", EXT_CASES, ids=[c[0] for c in EXT_CASES])
def test_fp64_extended_solver_parity(pf, name, pd, phi, xB, cap):
    """Extrapolated warm start + Anderson mixing: GPU FP64 against the oracle
    running the same scheme, 1e-10 and identical iteration counts."""
    o = _oracle(pd, phi, xB, cap)
    sim = _gpu(pf, pd, phi, xB, cap, pf.PF_FP64)
    for s in range(6):
        o.step()
        sim.step()
        sc = sim.scalars()
        dphi, dY, dx = _compare(o, sim, pf)
        print(f"{name} step {s + 1}: dphi={dphi:.2e} dY={dY:.2e} dxB={dx:.2e} "
              f"iters gpu={sc['last_elastic_iterations']} cpu={o.last_elastic_iterations}", flush=True)
        assert sc["last_elastic_iterations"] == o.last_elastic_iterations
        assert dphi < FP64_TOL and dY < FP64_TOL and dx < FP64_TOL


@pytest.mark.gpu
def test_mixed_extended_against_fp64_reference(pf):
    """Mixed precision with the accelerated solver against the plain FP64
    oracle (the reference iteration): x_B within the 1e-5 gate after 8 steps,
    mass conserved to rounding, fewer elastic updates."""
    pd, phi, xB, meta = config(2, 16)
    cap = meta["xB_max_safe"]
    o = _oracle(pd, phi, xB, cap)
    sim = _gpu(pf, _ext(pd), phi, xB, cap, pf.PF_MIXED)
    it_o = it_g = 0
    for s in range(8):
        o.step()
        sim.step()
        it_o += o.last_elastic_iterations
        it_g += sim.scalars()["last_elastic_iterations"]
    dphi, dY, dx = _compare(o, sim, pf)
    print(f"mixed+ext vs fp64 plain: dxB={dx:.2e} dphi={dphi:.2e} updates {it_g} vs {it_o}", flush=True)
    assert dx < MIXED_XB_TOL
    assert it_g < it_o
    sc = sim.scalars()
    assert abs(sc["mass_xBtot"] - o.target_mass) <= 1e-12 * max(abs(o.target_mass), 1.0)


@pytest.mark.gpu
@pytest.mark.parametrize("max_it", [1, 2, 24])
def test_zero_mode_iteration_cap(pf, max_it):
    """zero_mode_max_iterations caps the root find (main_cuda.cu:4360-4395):
    the GPU step fails with PF_E_ZERO_MODE exactly when the oracle's search
    stops short of the tolerance, and otherwise takes the same steps.  A cap
    of 0 is refused when the handle is created (main_cuda.cu:24468-24475;
    tests/test_abi.py)."""
    from oracle.step_ref import ZeroModeRejected
    name, pd, phi, xB, cap = CASES[2]
    pd = dict(pd, zero_mode_max_iterations=max_it)
    o = _oracle(pd, phi, xB, cap)
    sim = _gpu(pf, pd, phi, xB, cap, pf.PF_FP64)
    for s in range(3):
        try:
            o.step()
            oracle_ok = True
        except ZeroModeRejected:
            oracle_ok = False
        if oracle_ok:
            sim.step()
            sc = sim.scalars()
            print(f"max_it={max_it} step {s + 1}: iters gpu={sc['zero_mode_iterations']} "
                  f"cpu={o.zero_mode_iterations}", flush=True)
            assert sc["zero_mode_iterations"] == o.zero_mode_iterations
        else:
            with pytest.raises(pf.PfError) as ei:
                sim.step()
            print(f"max_it={max_it} step {s + 1}: both rejected ({ei.value})", flush=True)
            assert ei.value.code == -5 and "zero-mode" in str(ei.value)
            return
    assert max_it > 0


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_elastic_state_is_last_converged_solution(pf, prec):
    """After a step of the accelerated solver the elastic state is the
    displacement that step converged to (the history ring rotates buffers
    rather than copying them): a cold reference-iteration solve at the same
    phi and x_B reaches the same field to the solver tolerance."""
    pd, phi, xB, meta = config(2, 16)
    cap = meta["xB_max_safe"]
    a = _gpu(pf, _ext(pd), phi, xB, cap, pf.PF_FP64 if prec == "fp64" else pf.PF_MIXED)
    for _ in range(5):
        a.step()
    phi_n, xB_n = a.field(pf.PF_PHI), a.field(pf.PF_XB)
    a.step()   # its solve is at (phi_n, xB_n)
    ua = a.elastic_state()
    b = _gpu(pf, pd, phi_n, xB_n, float("inf"), pf.PF_FP64)
    b.step()
    ub = b.elastic_state()
    rel = np.abs(ua - ub).max() / np.abs(ub).max()
    print(f"{prec}: |u_accel - u_cold| / |u| = {rel:.2e}", flush=True)
    assert rel < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("name,pd,phi,xB,cap", [CASES[1], CASES[2]], ids=[CASES[1][0], CASES[2][0]])
def test_stress_components(pf, name, pd, phi, xB, cap):
    """PF_SIGMA_*: the step's stress at its start-of-step fields, against the
    oracle's sigma of the same step (FP64 to 1e-10 of the stress scale; the
    mixed path's FP32 stress to 1e-5 of it); unavailable until kept."""
    o = _oracle(pd, phi, xB, cap)
    sim = _gpu(pf, pd, phi, xB, cap, pf.PF_FP64)
    mixed = _gpu(pf, pd, phi, xB, cap, pf.PF_MIXED)
    with pytest.raises(pf.PfError, match="pf_keep_stress"):
        sim.field(pf.PF_SIGMA_XX)
    sim.keep_stress()
    mixed.keep_stress()
    for s in range(2):
        o.step()
        sim.step()
        mixed.step()
    scale = max(float(np.abs(c).max()) for c in o.sigma)
    assert scale > 0.0
    tr = np.zeros(o.shape)
    for i, which in enumerate(pf.PF_SIGMA):
        g = sim.field(which)
        err = float(np.abs(g - o.sigma[i]).max())
        errm = float(np.abs(mixed.field(which) - o.sigma[i]).max())
        print(f"{name} sigma[{i}]: fp64 {err / scale:.2e}  mixed {errm / scale:.2e} (of {scale:.3e})", flush=True)
        assert err <= 1e-10 * scale and errm <= 1e-5 * scale
        tr += g if i < 3 else 0.0
    assert np.allclose(tr, sim.field(pf.PF_TRSIGMA), rtol=0, atol=1e-12 * scale)
    sim.keep_stress(False)
    sim.step()
    with pytest.raises(pf.PfError, match="pf_keep_stress"):
        sim.field(pf.PF_SIGMA_XY)


SOLVERS = [("fp64_reference", "PF_FP64", {}), ("mixed", "PF_MIXED", {}),
           ("fp64_accelerated", "PF_FP64", {"elastic_warm_start_order": 3, "elastic_anderson_depth": 3})]


@pytest.mark.gpu
@pytest.mark.parametrize("name,prec,extra", SOLVERS, ids=[c[0] for c in SOLVERS])
def test_elastic_nonconvergence_is_fail_closed(pf, name, prec, extra):
    """An elastic solve that reaches elastic_iter_max above tolerance fails the
    step with PF_E_NONCONVERGED when elastic_fail_on_nonconvergence is set
    (main_cuda.cu:16273-16279), and is accepted at the cap otherwise."""
    pd, phi, xB, meta = config(2, 16)
    pd = dict(pd, elastic_iter_max=3, elastic_residual_tolerance=1e-15, **extra)
    s = _gpu(pf, dict(pd, elastic_fail_on_nonconvergence=1), phi, xB, meta["xB_max_safe"], getattr(pf, prec))
    with pytest.raises(pf.PfError, match="did not converge") as ei:
        s.step()
    assert ei.value.code == -4
    # the failed step left the handle part-way: it refuses to step until restored
    with pytest.raises(pf.PfError, match="holds no valid state") as ei:
        s.step()
    assert ei.value.code == -1
    s.set_fields(phi, xB, xB_cap=meta["xB_max_safe"])
    with pytest.raises(pf.PfError, match="did not converge"):
        s.step()
    t = _gpu(pf, dict(pd, elastic_fail_on_nonconvergence=0), phi, xB, meta["xB_max_safe"], getattr(pf, prec))
    t.step(2)
    sc = t.scalars()
    assert sc["last_elastic_iterations"] == 3 and sc["last_elastic_residual"] > 1e-15
    assert np.isfinite(t.field(pf.PF_XB)).all()


@pytest.mark.gpu
def test_nonfinite_initial_fields_are_refused(pf):
    """NaN/Inf in phi, x_B or dY/dt is refused with PF_E_NONFINITE before any
    device state changes (load_raw_init_fields, main_cuda.cu:2905-2965); the
    handle then accepts valid fields."""
    pd, phi, xB, meta = config(2, 16)
    s = pf.Simulation(pd, pf.PF_FP64)
    for bad_phi, bad_xB in ((True, False), (False, True)):
        p = phi.copy()
        x = xB.copy()
        if bad_phi:
            p[3, 4, 5] = np.nan
        if bad_xB:
            x[0, 0, 1] = np.inf
        with pytest.raises(pf.PfError, match="NaN/Inf") as ei:
            s.set_fields(p, x, xB_cap=meta["xB_max_safe"])
        assert ei.value.code == -6
    s.set_fields(phi, xB, xB_cap=meta["xB_max_safe"])
    s.step()
    assert np.isfinite(s.field(pf.PF_XB)).all()


@pytest.mark.gpu
def test_y_update_single_mode_recurrence(pf):
    """The HIP composition update on a small single Fourier mode of Y over a
    uniform matrix follows the closed recurrence of the lagged-RHS scheme,
    a_{n+1} (1 + dt Db k^2) = a_n + (1 - q)(a_n - a_{n-1}), Db = D q,
    q = x(1-x) (FP64 path; the oracle test states the derivation)."""
    shape = (16, 4, 4)
    pd = base_params(shape, dt=2.0, hetero=False, zero_mode=False, elastic=False)
    x0, eps, m = 0.03, 1e-7, 4
    Y0 = math.log(x0 / (1.0 - x0))
    cosx = np.cos(2.0 * np.pi * m * np.arange(shape[0]) / shape[0])
    Y = Y0 + eps * np.broadcast_to(cosx[:, None, None], shape)
    s = pf.Simulation(pd, pf.PF_FP64)
    s.set_fields(np.zeros(shape), 1.0 / (1.0 + np.exp(-Y)))

    def amp(f):
        return 2.0 * float(np.mean(f * cosx[:, None, None]))

    a = [amp(s.field(pf.PF_Y))]
    for _ in range(6):
        s.step()
        a.append(amp(s.field(pf.PF_Y)))
    q = x0 * (1.0 - x0)
    den = 1.0 + pd["dt"] * pd["D_alpha"] * q * (2.0 * np.pi * m / shape[0]) ** 2
    b = [a[0], a[0] / den]
    for n in range(1, 6):
        b.append((b[n] + (1.0 - q) * (b[n] - b[n - 1])) / den)
    err = max(abs(a[n] - b[n]) for n in range(7)) / abs(a[0])
    print(f"single-mode Y recurrence: max deviation {err:.2e} of the amplitude", flush=True)
    assert err <= 1e-6


@pytest.mark.gpu
def test_oracle_resumes_from_gpu_state(pf):
    """The CPU baseline of bench.py starts the oracle from the FP64 GPU state
    (fields, converged displacement expanded from the compact spectrum,
    frozen zero-mode target).  From that state the oracle's next step must
    equal the GPU's own next step: same elastic updates, fields to 1e-10."""
    import bench
    pd, phi, xB, meta = config(2, 16)
    cap = meta["xB_max_safe"]
    sim = _gpu(pf, pd, phi, xB, cap, pf.PF_FP64)
    sim.step(3)
    snap = bench.gpu_state_snapshot(pf, sim)
    o = Oracle(Params(**pd), snap["phi"], snap["xB"], dYdt0=snap["dYdt"], xB_cap=cap)
    o.Y, o.xB = snap["Y"].copy(), snap["xB"].copy()
    o.u_k = [c.copy() for c in snap["u_k"]]
    o.warm_valid = True
    o.target_mass = snap["target_mass"]
    o.step()
    sim.step()
    dphi, dY, dx = _compare(o, sim, pf)
    print(f"resume: dphi={dphi:.2e} dY={dY:.2e} iters gpu={sim.scalars()['last_elastic_iterations']} "
          f"cpu={o.last_elastic_iterations}", flush=True)
    assert sim.scalars()["last_elastic_iterations"] == o.last_elastic_iterations
    assert dphi < FP64_TOL and dY < FP64_TOL and dx < FP64_TOL
