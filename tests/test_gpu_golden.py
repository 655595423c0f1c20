"""GPU FP64 path against the committed oracle fixtures (tests/golden/).

cfg1_64_100: workload 1 at its full size (64^3 sphere, homogeneous elasticity,
zero mode, dt 0.01) for 100 steps, checked on three central planes and
whole-field checksums to the FP64 gate of 1e-10; mass conserved to rounding;
energy densities of the last step.
cfg2_16_10: the heterogeneous workload geometry at 16^3, 10 steps, full fields
and the per-step elastic iteration counts."""
import os

import numpy as np
import pytest

from cuda_sto_pf_amd.configs import config

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-10


@pytest.mark.gpu
def test_cfg1_64_100_steps(pf):
    g = np.load(os.path.join(HERE, "cfg1_64_100.npz"))
    pd, phi, xB, meta = config(1)
    sim = pf.Simulation(pd, pf.PF_FP64)
    sim.set_fields(phi, xB, xB_cap=meta["xB_max_safe"])
    iters = []
    for s in range(int(g["steps"])):
        sim.step()
        iters.append(sim.scalars()["last_elastic_iterations"])
    assert iters == list(g["iters"])
    worst = 0.0
    for name, fid in (("phi", pf.PF_PHI), ("Y", pf.PF_Y), ("xB", pf.PF_XB)):
        f = sim.field(fid)
        n = f.shape
        for tag, plane in (("px", f[n[0] // 2]), ("py", f[:, n[1] // 2]), ("pz", f[:, :, n[2] // 2])):
            d = np.abs(plane - g[f"{name}_{tag}"]).max()
            worst = max(worst, d)
            assert d < TOL, (name, tag, d)
        sums = np.array([f.sum(), (f * f).sum(), f.min(), f.max()])
        ref = g[f"{name}_sums"]
        assert np.all(np.abs(sums - ref) <= TOL * np.maximum(1.0, np.abs(ref)) * f.size ** 0.5), (name, sums, ref)
    sc = sim.scalars()
    assert abs(sc["mass_xBtot"] - float(g["target_mass"])) <= 1e-12 * abs(float(g["target_mass"]))
    e = np.array([sc["f_grad"], sc["f_dw"], sc["f_elastic"], sc["f_chem"]])
    assert np.allclose(e, g["energies"], rtol=1e-10, atol=1e-14)
    print(f"cfg1 64^3 x100: worst plane |gpu - oracle| = {worst:.3e}", flush=True)


@pytest.mark.gpu
def test_cfg2_16_10_steps(pf):
    g = np.load(os.path.join(HERE, "cfg2_16_10.npz"))
    pd, phi, xB, meta = config(2, 16)
    sim = pf.Simulation(pd, pf.PF_FP64)
    sim.set_fields(phi, xB, xB_cap=meta["xB_max_safe"])
    iters = []
    for s in range(10):
        sim.step()
        iters.append(sim.scalars()["last_elastic_iterations"])
    assert iters == list(g["iters"])
    for name, fid in (("phi", pf.PF_PHI), ("Y", pf.PF_Y), ("xB", pf.PF_XB), ("dYdt", pf.PF_DYDT)):
        d = np.abs(sim.field(fid) - g[name]).max()
        assert d < TOL * (1e2 if name == "dYdt" else 1.0), (name, d)


@pytest.mark.gpu
def test_cfg2_32_100_steps(pf):
    """Heterogeneous elasticity over 100 steps at 32^3 (workload 2 geometry):
    the full phi and Y fields within the FP64 gate of 1e-10, identical
    elastic iteration counts every step, x_B checksums, mass to rounding and
    the energy densities of the last step."""
    g = np.load(os.path.join(HERE, "cfg2_32_100.npz"))
    pd, phi, xB, meta = config(2, 32)
    sim = pf.Simulation(pd, pf.PF_FP64)
    sim.set_fields(phi, xB, xB_cap=meta["xB_max_safe"])
    iters = []
    for s in range(100):
        sim.step()
        iters.append(sim.scalars()["last_elastic_iterations"])
    assert iters == list(g["iters"])
    dphi = np.abs(sim.field(pf.PF_PHI) - g["phi"]).max()
    dY = np.abs(sim.field(pf.PF_Y) - g["Y"]).max()
    print(f"cfg2 32^3 x100: max|dphi|={dphi:.3e} max|dY|={dY:.3e}", flush=True)
    assert dphi < TOL and dY < TOL
    x = sim.field(pf.PF_XB)
    sums = np.array([x.sum(), (x * x).sum(), x.min(), x.max()])
    assert np.all(np.abs(sums - g["xB_sums"]) <= TOL * np.maximum(1.0, np.abs(g["xB_sums"])) * 8), (sums, g["xB_sums"])
    sc = sim.scalars()
    assert abs(sc["mass_xBtot"] - float(g["target_mass"])) <= 1e-12 * abs(float(g["target_mass"]))
    e = np.array([sc["f_grad"], sc["f_dw"], sc["f_elastic"], sc["f_chem"]])
    assert np.allclose(e, g["energies"], rtol=1e-10, atol=1e-14)
