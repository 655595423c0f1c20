"""Opt-in implementation variants selected by environment at library load:
PF_ZPASS=reg (the register-resident z pass, pf_regfft.hpp) and PF_KZ_PAD=0
(spectral rows not padded to whole cache lines).  Each runs in a child
process (the selection is read once per process) and must reproduce the FP64
oracle step to the same 1e-10 as the default path."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from cuda_sto_pf_amd import pfcore as pf
from cuda_sto_pf_amd.configs import base_params, composition, spheres_phi
from oracle.step_ref import Oracle, Params
shape = tuple(json.loads(sys.argv[2]))
pd = base_params(shape, hetero=True, zero_mode=True, elastic=True)
rng = np.random.default_rng(11)
phi = spheres_phi(shape, [(shape[0] / 2, shape[1] / 2, shape[2] / 2)], [3.0], width=1.5)
xB = composition(phi, rng)
o = Oracle(Params(**pd), phi, xB, xB_cap=0.5)
s = pf.Simulation(pd, pf.PF_FP64)
s.set_fields(phi, xB, xB_cap=0.5)
worst = 0.0
for _ in range(3):
    o.step()
    s.step()
    for f, ref in ((pf.PF_PHI, o.phi), (pf.PF_Y, o.Y), (pf.PF_XB, o.xB)):
        worst = max(worst, float(np.abs(s.field(f) - ref).max()))
print(json.dumps({"worst": worst, "iters": s.scalars()["last_elastic_iterations"],
                  "oracle_iters": o.last_elastic_iterations}))
"""

CASES = [("zpass_reg_32", {"PF_ZPASS": "reg"}, (32, 32, 32)),
         ("zpass_reg_kzpad", {"PF_ZPASS": "reg"}, (24, 20, 128)),
         ("no_kz_padding", {"PF_KZ_PAD": "0"}, (24, 20, 128))]


@pytest.mark.gpu
@pytest.mark.parametrize("name,env,shape", CASES, ids=[c[0] for c in CASES])
def test_variant_matches_oracle(name, env, shape):
    print(f"variant {name} {env} shape={shape}", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT, json.dumps(shape)], capture_output=True, text=True,
                       timeout=240, env=dict(os.environ, **env))
    print(r.stdout[-800:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["worst"] < 1e-10
    assert res["iters"] == res["oracle_iters"]
