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


KZ_CHILD = r"""
import hashlib, json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from cuda_sto_pf_amd import pfcore as pf
from cuda_sto_pf_amd.configs import base_params, composition, spheres_phi
shape = (24, 20, 128)
pd = dict(base_params(shape, hetero=True, zero_mode=True, elastic=True), elastic_warm_start_order=3,
          elastic_anderson_depth=3)
rng = np.random.default_rng(12)
phi = spheres_phi(shape, [(12.0, 10.0, 64.0)], [4.0], width=1.5)
xB = composition(phi, rng)
out = {}
for name, prec in (("fp64", pf.PF_FP64), ("mixed", pf.PF_MIXED)):
    s = pf.Simulation(pd, prec)
    s.set_fields(phi, xB, xB_cap=0.5)
    s.step(4)
    h = hashlib.sha256()
    for f in (pf.PF_PHI, pf.PF_Y, pf.PF_XB):
        h.update(s.field(f).tobytes())
    out[name] = {"hash": h.hexdigest(), "kz_stride": s.scalars()["kz_stride"]}
print(json.dumps(out))
"""


@pytest.mark.gpu
def test_kz_stride_choice_is_bitwise_neutral():
    """The plan tuner's kz strides (whole lines, half lines, none) move only
    where spectra sit: four steps give bitwise the same fields with each
    forced stride and with the tuned one."""
    res = {}
    for pad in ("16", "4", "0", None):
        env = dict(os.environ)
        env.pop("PF_KZ_PAD", None)
        if pad is not None:
            env["PF_KZ_PAD"] = pad
        print(f"kz stride child PF_KZ_PAD={pad}", flush=True)
        r = subprocess.run([sys.executable, "-c", KZ_CHILD, ROOT], capture_output=True, text=True, timeout=240,
                           env=env)
        print(r.stdout[-600:], r.stderr[-1200:], flush=True)
        assert r.returncode == 0
        res[pad] = json.loads(r.stdout.strip().splitlines()[-1])
    for prec in ("fp64", "mixed"):
        strides = {res[p][prec]["kz_stride"] for p in ("16", "4", "0")}
        assert len(strides) >= 2, strides          # the forced strides differ
        assert len({res[p][prec]["hash"] for p in res}) == 1, prec
        assert res[None][prec]["kz_stride"] in {48, 44, 43}
