"""GP-zone eta update (kernel-level parity): pf_eta_update against the FP64
oracle (oracle/eta_ref.py) on seeded random fields, chemical and raw-reaction
drives, with and without the elastic coupling, cubic and rectangular grids."""

import numpy as np
import pytest

from cuda_sto_pf_amd.configs import base_params
from oracle.eta_ref import EtaParams, eta_update
from oracle.step_ref import Params

CASES = [
    ("chem_16", (16, 16, 16), dict(gp_W_eta=0.8, gp_kappa_eta=1.5, gp_L_eta=2.0, gp_delta_g0=0.02)),
    ("raw_elastic_rect", (12, 16, 8), dict(gp_W_eta=0.3, gp_kappa_eta=2.0, gp_L_eta=1.0, gp_raw_reaction_drive_only=1,
                                           gp_mu_reference_raw=-2.5e4, gp_elastic=1, gp_eps_iso=0.012,
                                           gp_elastic_derivative_scale=0.7)),
    ("raw_units_20", (20, 20, 20), dict(gp_W_eta=0.1, gp_kappa_eta=0.5, gp_L_eta=0.5, gp_raw_reaction_drive_only=1,
                                        gp_raw_reaction_drive_use_raw_units_debug=1, gp_mu_reference_raw=-2.9e4)),
]


def _fields(shape, seed):
    rng = np.random.default_rng(seed)
    x = np.arange(shape[0])[:, None, None]
    eta = np.clip(0.5 + 0.4 * np.sin(2 * np.pi * x / shape[0]) + 0.05 * rng.standard_normal(shape), 0, 1)
    phi = rng.uniform(0.0, 0.7, shape)
    xa = rng.uniform(0.002, 0.08, shape)
    trs = rng.normal(0.0, 0.3, shape)
    return eta, phi, xa, trs


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("name,shape,q", CASES, ids=[c[0] for c in CASES])
def test_eta_update_parity(pf, name, shape, q, prec):
    pd = base_params(shape, elastic=False, zero_mode=False)
    eta, phi, xa, trs = _fields(shape, 7)
    ref = eta_update(Params(**pd), EtaParams(**q), eta, phi, xa, trs)
    sim = pf.Simulation(pd, pf.PF_FP64 if prec == "fp64" else pf.PF_MIXED)
    got = sim.eta_update(q, eta, phi, xa, trs)
    err = np.abs(got - ref).max()
    print(f"eta {name} {prec}: max|d|={err:.2e} range [{ref.min():.3f}, {ref.max():.3f}]", flush=True)
    assert err <= (1e-12 if prec == "fp64" else 1e-5)
    assert set(EtaParams.__dataclass_fields__) >= set(q)
